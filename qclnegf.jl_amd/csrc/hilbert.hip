// FFT Hilbert transform of every column alpha = (m,a,b) of a field
// (fft_parallel.jl:27-291; production_operators.jl:735-884).
//
// Lambda_e = sum_{e' != e} wE_e' Gamma_e' / (2 pi (e - e') de) is evaluated as
// the zero-padded linear convolution of wE*Gamma with the kernel
// 1/(2 pi lag de), lag = -(NE-1)..NE-1, through length-L FFTs
// (L = nextpow2(2NE-1)), extracting y[e + NE - 1]. The round-off witness
// repeats the transform with 2L and reports ||L2 - L|| / (||L|| + 1e-14).
//
// One workgroup per column. Each thread keeps 16 complex values in registers
// (T = L/16 threads); every Stockham pass does radix-R butterflies in
// registers and exchanges data through LDS one real component at a time, so
// the LDS footprint is L doubles.
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
    int radix[2][16] = {};  // pass radices for L and 2L
    int npass[2] = {0, 0};
    qclhip::c64* twiddle = nullptr;  // exp(-2 pi i idx / Lmax), idx < Lmax
    qclhip::c64* kf[2] = {nullptr, nullptr};  // kernel spectra for L and 2L
    int Lmax = 0;
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

template <int V, int R, bool INV>
__device__ inline void butterflies(double* re, double* im, int t, int T, int L, int Ns,
                                   const c64* __restrict__ tw, int twstride) {
#pragma unroll
    for (int i = 0; i < V / R; ++i) {
        const int j = t + i * T;
        if (j >= L / R) break;
        const int k = j % Ns;
        if (Ns > 1) {
            const int step = (L / (Ns * R)) * twstride;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                c64 w = tw[r * k * step];
                const double wi = INV ? -w.y : w.y;
                const double xr = re[i * R + r], xi = im[i * R + r];
                re[i * R + r] = w.x * xr - wi * xi;
                im[i * R + r] = w.x * xi + wi * xr;
            }
        }
        dft<R, INV>(re + i * R, im + i * R);
    }
}

template <int V>
__device__ inline void run_butterflies(int R, bool inv, double* re, double* im, int t, int T, int L,
                                       int Ns, const c64* tw, int twstride) {
    if (inv) {
        switch (R) {
            case 16: butterflies<V, 16, true>(re, im, t, T, L, Ns, tw, twstride); break;
            case 8: butterflies<V, 8, true>(re, im, t, T, L, Ns, tw, twstride); break;
            case 4: butterflies<V, 4, true>(re, im, t, T, L, Ns, tw, twstride); break;
            default: butterflies<V, 2, true>(re, im, t, T, L, Ns, tw, twstride); break;
        }
    } else {
        switch (R) {
            case 16: butterflies<V, 16, false>(re, im, t, T, L, Ns, tw, twstride); break;
            case 8: butterflies<V, 8, false>(re, im, t, T, L, Ns, tw, twstride); break;
            case 4: butterflies<V, 4, false>(re, im, t, T, L, Ns, tw, twstride); break;
            default: butterflies<V, 2, false>(re, im, t, T, L, Ns, tw, twstride); break;
        }
    }
}

// Index of value (i, r) after a pass of radix R with span Ns.
__device__ inline int out_index(int j, int r, int R, int Ns) { return (j / Ns) * Ns * R + (j % Ns) + r * Ns; }

// Full-length transform. On entry the registers hold the inputs of the first
// pass (value (i,r) = x[j + r*L/R0]); on exit the outputs of the last pass
// (value (i,r) at index j + r*L/Rlast).
template <int V>
__device__ void fft_passes(double* re, double* im, double* lds, int t, int T, int L, const int* radix,
                           int npass, bool inv, const c64* tw, int twstride) {
    int Ns = 1;
    for (int p = 0; p < npass; ++p) {
        const int R = inv ? radix[npass - 1 - p] : radix[p];
        run_butterflies<V>(R, inv, re, im, t, T, L, Ns, tw, twstride);
        if (p == npass - 1) break;
        const int Rn = inv ? radix[npass - 2 - p] : radix[p + 1];
        // Register index f = i*R + r is static in every unrolled loop below.
        for (int comp = 0; comp < 2; ++comp) {
            double* v = comp ? im : re;
            __syncthreads();
#pragma unroll
            for (int f = 0; f < V; ++f) {
                const int i = f / R, r = f % R;
                const int j = t + i * T;
                if (j < L / R) lds[out_index(j, r, R, Ns)] = v[f];
            }
            __syncthreads();
#pragma unroll
            for (int f = 0; f < V; ++f) {
                const int i = f / Rn, r = f % Rn;
                const int j = t + i * T;
                if (j < L / Rn) v[f] = lds[j + r * (L / Rn)];
            }
        }
        Ns *= R;
    }
}

struct HilbertArgs {
    int NE, L, npass;
    int radix[16];
    long long ncols;
    const double* wE;
    const c64* gamma;   // Gamma input, or NULL when (SL, SG) are given
    const c64* SL;
    const c64* SG;
    c64* out;           // Lambda (mode 0) or S^R = Lambda - i Gamma/2 (mode 1)
    c64* lambda_copy;   // optional copy of Lambda for the witness
    const c64* lambda_ref;  // witness mode: Lambda of the L transform
    double* partial;    // witness mode: [2][ncols] sums |L2-L|^2 and |L|^2
    const c64* tw;
    int twstride;
    const c64* kf;
    int mode;           // 0: Lambda, 1: S^R, 2: witness only
};

__device__ inline c64 load_gamma(const HilbertArgs& a, size_t idx) {
    if (a.gamma) return a.gamma[idx];
    c64 l = a.SL[idx], g = a.SG[idx];
    // Gamma = i (S> - S<)
    return cmk(-(g.y - l.y), g.x - l.x);
}

template <int V>
__global__ __launch_bounds__(512) void k_hilbert(HilbertArgs a) {
    extern __shared__ __align__(16) double lds[];
    const int t = threadIdx.x, T = blockDim.x;
    const long long col = blockIdx.x;
    const int L = a.L, NE = a.NE;
    const size_t base = (size_t)col * NE;
    double re[V], im[V];
    // Load wE*Gamma, zero padded, into the first pass's input order.
    {
        const int R0 = a.radix[0];
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R0, r = f % R0;
            const int j = t + i * T;
            const int e = j + r * (L / R0);
            double xr = 0.0, xi = 0.0;
            if (j < L / R0 && e < NE) {
                c64 gm = load_gamma(a, base + e);
                xr = a.wE[e] * gm.x;
                xi = a.wE[e] * gm.y;
            }
            re[f] = xr;
            im[f] = xi;
        }
    }
    fft_passes<V>(re, im, lds, t, T, L, a.radix, a.npass, false, a.tw, a.twstride);
    // Multiply by the kernel spectrum (outputs at j + r*L/Rlast).
    const int Rl = a.radix[a.npass - 1];
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / Rl, r = f % Rl;
        const int j = t + i * T;
        if (j < L / Rl) {
            c64 k = a.kf[j + r * (L / Rl)];
            const double xr = re[f], xi = im[f];
            re[f] = k.x * xr - k.y * xi;
            im[f] = k.x * xi + k.y * xr;
        }
    }
    // Inverse transform (radices in reverse order, so no exchange is needed).
    fft_passes<V>(re, im, lds, t, T, L, a.radix, a.npass, true, a.tw, a.twstride);
    const double invL = 1.0 / (double)L;
    const int Rf = a.radix[0];
    double dn = 0.0, nn = 0.0;
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / Rf, r = f % Rf;
        const int j = t + i * T;
        const int idx = j + r * (L / Rf);
        const int e = idx - (NE - 1);
        if (j >= L / Rf || e < 0 || e >= NE) continue;
        c64 lam = cmk(re[f] * invL, im[f] * invL);
        const size_t gi = base + e;
        if (a.mode == 2) {
            c64 ref = a.lambda_ref[gi];
            const double dx = lam.x - ref.x, dy = lam.y - ref.y;
            dn += dx * dx + dy * dy;
            nn += ref.x * ref.x + ref.y * ref.y;
            continue;
        }
        if (a.lambda_copy) a.lambda_copy[gi] = lam;
        if (a.mode == 1) {
            c64 gm = load_gamma(a, gi);
            // S^R = Lambda - 0.5 i Gamma
            a.out[gi] = cmk(lam.x + 0.5 * gm.y, lam.y - 0.5 * gm.x);
        } else {
            a.out[gi] = lam;
        }
    }
    if (a.mode != 2) return;
    // Deterministic block reduction of the witness sums (fixed tree in LDS;
    // the transform no longer needs it).
    double* rd = lds;
    double* rn = lds + T;
    __syncthreads();
    rd[t] = dn;
    rn[t] = nn;
    __syncthreads();
    for (int width = 256; width > 0; width >>= 1) {
        if (t < width && t + width < T) {
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

__global__ __launch_bounds__(1024) void k_witness_reduce(const double* partial, long long ncols, double* out) {
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
    if (t == 0) out[0] = sqrt(red[0][0]) / (sqrt(red[1][0]) + 1e-14);
}

// Host: radices for a power-of-two length (16s first, remainder last).
int plan_radices(int L, int* radix) {
    int n = 0, rest = L;
    while (rest >= 16 && (rest % 16) == 0) {
        radix[n++] = 16;
        rest /= 16;
    }
    if (rest > 1) radix[n++] = rest;  // 2, 4 or 8
    return n;
}

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

std::vector<c64> kernel_spectrum(int NE, int L, double de) {
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<std::complex<long double>> k(L, 0.0L);
    for (int lag = -(NE - 1); lag <= NE - 1; ++lag) {
        if (lag == 0) continue;
        // The reference builds the kernel in Float64: inv(2 pi lag de).
        double v = 1.0 / (2.0 * M_PI * (double)lag * de);
        k[lag + NE - 1] = (long double)v;
    }
    host_fft(k);
    std::vector<c64> out(L);
    for (int i = 0; i < L; ++i) out[i] = cmk((double)k[i].real(), (double)k[i].imag());
    (void)pi;
    return out;
}

}  // namespace

int hilbert_launch(qclhip_ctx* ctx, const qclhip_hilbert* p, const HilbertArgs& a_in, int which) {
    HilbertArgs a = a_in;
    a.L = which ? 2 * p->L : p->L;
    a.npass = p->npass[which];
    for (int i = 0; i < 16; ++i) a.radix[i] = p->radix[which][i];
    a.kf = p->kf[which];
    a.tw = p->twiddle;
    a.twstride = p->Lmax / a.L;
    const int Vv = a.L <= 8192 ? 16 : 32;
    const int T = std::max(1, a.L / Vv);
    const size_t lds = (size_t)std::max(a.L, 2 * T) * sizeof(double);
    if (lds > ctx->max_lds_per_block || T > 512)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "FFT length exceeds the LDS of one workgroup");
    const void* fn = Vv == 16 ? (const void*)k_hilbert<16> : (const void*)k_hilbert<32>;
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (Vv == 16)
        hipLaunchKernelGGL(k_hilbert<16>, dim3((unsigned)p->ncols), dim3(T), lds, ctx->stream, a);
    else
        hipLaunchKernelGGL(k_hilbert<32>, dim3((unsigned)p->ncols), dim3(T), lds, ctx->stream, a);
    return launch_check(ctx, "k_hilbert");
}

int hilbert_run(qclhip_ctx* ctx, const qclhip_hilbert* p, const double* Gamma, const double* SL,
                const double* SG, double* out, int mode, double* roundoff_host) {
    const long long n = (long long)p->NE * p->ncols;
    HilbertArgs a{};
    a.NE = p->NE;
    a.ncols = p->ncols;
    a.wE = p->grid->wE;
    a.gamma = (const c64*)Gamma;
    a.SL = (const c64*)SL;
    a.SG = (const c64*)SG;
    a.out = (c64*)out;
    a.mode = mode;
    c64* lam = nullptr;
    double* partial = nullptr;
    if (roundoff_host) {
        if (!p->verify) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the FFT plan was built without roundoff verification");
        char* ws = (char*)workspace(ctx, (size_t)n * sizeof(c64) + (size_t)2 * p->ncols * sizeof(double) + 64);
        if (!ws) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
        lam = (c64*)ws;
        partial = (double*)(ws + (size_t)n * sizeof(c64));
        a.lambda_copy = mode == 0 ? nullptr : lam;
    }
    int rc = hilbert_launch(ctx, p, a, 0);
    if (rc || !roundoff_host) return rc;
    HilbertArgs w = a;
    w.mode = 2;
    w.out = nullptr;
    w.lambda_copy = nullptr;
    w.lambda_ref = mode == 0 ? (const c64*)out : lam;
    w.partial = partial;
    rc = hilbert_launch(ctx, p, w, 1);
    if (rc) return rc;
    double* res = partial + 2 * p->ncols;
    hipLaunchKernelGGL(k_witness_reduce, dim3(1), dim3(1024), 0, ctx->stream, partial, p->ncols, res);
    rc = launch_check(ctx, "k_witness_reduce");
    if (rc) return rc;
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
    const int Lmax = verify ? 2 * L : L;
    if ((size_t)Lmax * sizeof(double) > ctx->max_lds_per_block || Lmax > 16384)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "FFT length exceeds the LDS of one workgroup");
    auto* p = new qclhip_hilbert();
    p->grid = grid;
    p->NE = grid->NE;
    p->ncols = ncols;
    p->L = L;
    p->Lmax = Lmax;
    p->verify = verify ? 1 : 0;
    p->npass[0] = plan_radices(L, p->radix[0]);
    p->npass[1] = plan_radices(2 * L, p->radix[1]);
    // Twiddles exp(-2 pi i idx/Lmax) in long double, rounded once.
    std::vector<c64> tw(Lmax);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < Lmax; ++i) {
        long double ang = -2.0L * pi * (long double)i / (long double)Lmax;
        tw[i] = cmk((double)cosl(ang), (double)sinl(ang));
    }
    bool ok = hipMalloc(&p->twiddle, (size_t)Lmax * sizeof(c64)) == hipSuccess &&
              hipMemcpy(p->twiddle, tw.data(), (size_t)Lmax * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
    for (int w = 0; w < (verify ? 2 : 1) && ok; ++w) {
        const int Lw = w ? 2 * L : L;
        std::vector<c64> kf = kernel_spectrum(grid->NE, Lw, grid->de);
        ok = hipMalloc(&p->kf[w], (size_t)Lw * sizeof(c64)) == hipSuccess &&
             hipMemcpy(p->kf[w], kf.data(), (size_t)Lw * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
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
    delete p;
    return QCLHIP_OK;
}

int qclhip_fft_hilbert(qclhip_ctx* ctx, const qclhip_hilbert* plan, const double* Gamma, double* Lambda,
                       double* roundoff_host) {
    if (!ctx || !plan || !Gamma || !Lambda) return QCLHIP_ERR_ARGUMENT;
    if (Gamma == Lambda) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    return hilbert_run(ctx, plan, Gamma, nullptr, nullptr, Lambda, 0, roundoff_host);
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
