// Shared definitions of the MI355X (gfx950) SCBA operator library.
//
// Data layout in HBM: every field is the reference's Array{ComplexF64,4} of
// shape (NE,Nk,Nb,Nb) in column-major order (energy fastest), interleaved
// (re,im) doubles. Element (e,m,a,b) sits at e + NE*(m + Nk*(a + Nb*b)), so a
// block (e,m) is a strided Nb x Nb matrix with element stride NE*Nk and
// consecutive energies of one element are contiguous in memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/qclhip.h"

namespace qclhip {

using c64 = double2;  // one complex double, (x,y) = (re,im)

__host__ __device__ inline c64 cmk(double re, double im) { c64 z; z.x = re; z.y = im; return z; }
__device__ inline c64 cadd(c64 a, c64 b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ inline c64 csub(c64 a, c64 b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ inline c64 cscale(double s, c64 a) { return cmk(s * a.x, s * a.y); }
__device__ inline c64 cmul(c64 a, c64 b) {
    return cmk(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// c + a*b
__device__ inline c64 cfma(c64 c, c64 a, c64 b) {
    return cmk(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
// c - a*b
__device__ inline c64 cfms(c64 c, c64 a, c64 b) {
    return cmk(fma(-a.x, b.x, fma(a.y, b.y, c.x)), fma(-a.x, b.y, fma(-a.y, b.x, c.y)));
}
// c + a*conj(b)
__device__ inline c64 cfmac(c64 c, c64 a, c64 b) {
    return cmk(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.y, b.x, fma(-a.x, b.y, c.y)));
}
__device__ inline c64 cconj(c64 a) { return cmk(a.x, -a.y); }
__device__ inline double cabs1(c64 a) { return fabs(a.x) + fabs(a.y); }
// Linear mixing (1 - alpha) o + alpha c (beta = 1 - alpha) with one explicit
// rounding order, so that every mixing kernel produces the same bits.
__device__ inline c64 mix1(c64 o, c64 c, double alpha, double beta) {
    return cmk(fma(alpha, c.x, beta * o.x), fma(alpha, c.y, beta * o.y));
}
// Calls f(std::integral_constant<int, I>) for I = First..Last-1: the index
// is a compile-time constant in the body, so a comparison with a runtime
// value never turns a register array access into a dynamic index.
template <int First, int Last, typename F>
__device__ inline void static_for(F&& f) {
    if constexpr (First < Last) {
        f(std::integral_constant<int, First>{});
        static_for<First + 1, Last>(f);
    }
}
// 1/z by Smith's algorithm (no overflow for |z| near the double range).
__device__ inline c64 crecip(c64 z) {
    if (fabs(z.x) >= fabs(z.y)) {
        double r = z.y / z.x, d = z.x + z.y * r;
        return cmk(1.0 / d, -r / d);
    }
    double r = z.x / z.y, d = z.y + z.x * r;
    return cmk(r / d, -1.0 / d);
}
// 1/a from the hardware reciprocal refined by two Newton steps (within an
// ulp or two for normal a; no IEEE division sequence).
__device__ inline double rcp_nr(double a) {
    double x = __builtin_amdgcn_rcp(a);
    double e = fma(-a, x, 1.0);
    x = fma(x, e, x);
    e = fma(-a, x, 1.0);
    return fma(x, e, x);
}
// 1/sqrt(a) for a normal positive a: the hardware estimate refined by two
// Newton steps.
__device__ inline double rsqrt_nr(double a) {
    double y = __builtin_amdgcn_rsq(a);
    double h = 0.5 * a * y;
    y = fma(y, fma(-h, y, 0.5), y);
    h = 0.5 * a * y;
    return fma(y, fma(-h, y, 0.5), y);
}
// The same quotient with selects instead of branches and two refined
// reciprocals (1/a and 1/d): for wave-uniform pivots of normal magnitude.
__device__ inline c64 crecip_sel(c64 z) {
    const bool sw = fabs(z.y) > fabs(z.x);
    const double a = sw ? z.y : z.x, b = sw ? z.x : z.y;
    const double r = b * rcp_nr(a);
    const double t = rcp_nr(fma(b, r, a));
    const double u = r * t;
    return sw ? cmk(u, -t) : cmk(t, -u);
}

// Device-side status word shared by all kernels of one context:
// code (QCLHIP_* error, 0 = fine) and the smallest one-based failing block.
struct DeviceStatus {
    int code;
    int pad;
    unsigned long long block;
};

// Paired-normalization coefficients produced on the device
// (reference/greens.jl:155-185) and read by the normalization kernel.
struct PairCoefficients {
    double a, c, lambda, number, holes;
};

__device__ inline void report_status(DeviceStatus* st, int code, unsigned long long block) {
    atomicCAS(&st->code, 0, code);
    if (block) atomicMin(&st->block, block);
}

// Wave-local LDS ordering: LDS instructions of one wavefront execute in
// order, so a compiler fence around a wave barrier makes one lane's LDS store
// visible to the other lanes' later loads.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }
__device__ inline int shfl_xor_i(int v, int mask) { return __shfl_xor(v, mask, 64); }
__device__ inline double readlane_d(double v, int lane) {
    long long bits = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
    int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

}  // namespace qclhip

// ---------------------------------------------------------------------------
// Host-side context (opaque to callers).
struct qclhip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    qclhip::DeviceStatus* status = nullptr;         // device
    qclhip::DeviceStatus* status_host = nullptr;    // pinned
    qclhip::PairCoefficients* coef = nullptr;       // device
    double* host_scalars = nullptr;                 // pinned, 64 doubles
    void* workspace = nullptr;                      // device scratch, grows on demand
    size_t workspace_bytes = 0;
    // Incremented whenever the scratch is reallocated: a captured graph holds
    // the scratch address of its capture and must not replay after a change.
    unsigned long long workspace_gen = 0;
    int sm_count = 0;
    size_t max_lds_per_block = 65536;
    std::string last_error;
    // Deferred results: entries that return host scalars write them to
    // defer_dev[defer_next...] on the device instead and do not synchronize.
    double* defer_dev = nullptr;
    int defer_cap = 0;
    int defer_next = 0;
    bool capturing = false;  // inside qclhip_graph_begin/_end
    // Autotune cache: (Nb, NE, Nk) -> chosen Keldysh engine and the two
    // measured times (ms).
    struct TuneEntry {
        int Nb, NE, Nk, engine, dyson_waves, keldysh_waves, keldysh_split;
        float ms_vector, ms_matrix;
        float ms_dyson[2], ms_keldysh[2];  // Dyson W = 4, 8; Keldysh W = 2, 4 (0: not possible)
        float ms_fused;                    // the fused Dyson + Keldysh kernel
        float ms_split;                    // the vector Keldysh kernel at the chosen waves per block
    };
    std::vector<TuneEntry> tune_cache;
};

struct qclhip_grid {
    int NE = 0, Nk = 0, Nb = 0, gs = 2;
    double de = 0.0;        // uniform spacing (0 when the grid is not uniform)
    double* eps = nullptr;  // device [NE]
    double* wE = nullptr;   // device [NE]
    double* wk = nullptr;   // device [Nk]
    int* flux_order = nullptr;  // device [NE] energies sorted by (|e|, e)
    std::vector<double> eps_host, wE_host, wk_host;
    // GF-solve precision and engines (qclhip_grid_set_precision / _set_engine).
    int precision = QCLHIP_PRECISION_FP64;
    int refine_steps = 2;
    double refine_tol = 1e-10;
    int keldysh_engine = QCLHIP_ENGINE_VECTOR;
    unsigned* block_flags = nullptr;  // device, one per (e,m) block (mixed precision)
    // Energies per workgroup of the FP64 vector Dyson / Keldysh kernels (0:
    // built-in rule), set by qclhip_grid_autotune.
    int dyson_waves = 0, keldysh_waves = 0;
    int keldysh_split = 0;  // waves per block of the FP64 vector Keldysh kernel (0: the rule)
};

struct qclhip_shift {
    int NE = 0;
    double delta = 0.0;
    int* left = nullptr;
    int* right = nullptr;
    double* wr = nullptr;
    unsigned char* valid = nullptr;
    // Transpose of the plan (CSR over source nodes): tptr[NE+1], tsrc, tw.
    int* tptr = nullptr;
    int* tsrc = nullptr;
    double* tw = nullptr;
};

namespace qclhip {

// Fixed device scratch of the allocation-free entry points (16 MiB per device).
constexpr size_t kStaticScratchBytes = (size_t)16 << 20;
// Looks up the device's copy of that scratch once, with `device` current.
int static_scratch_setup(int device);
// Its device address (nullptr before set-up).
void* static_scratch(int device);

int fail(qclhip_ctx* ctx, int code, const std::string& msg);
int hip_fail(qclhip_ctx* ctx, hipError_t err, const char* what);
// Device scratch of at least `bytes`, owned by the context.
void* workspace(qclhip_ctx* ctx, size_t bytes);
int launch_check(qclhip_ctx* ctx, const char* what);
// In deferred-results mode, copy n device doubles to the next result slots
// and set *deferred; otherwise leave *deferred false.
int defer_result(qclhip_ctx* ctx, const double* src, int n, bool* deferred);
// The same for the paired coefficients, in out_host order
// (lambda, number, holes, a, c).
int defer_coefficients(qclhip_ctx* ctx, const PairCoefficients* src, bool* deferred);
// Gamma = i (S> - S<) elementwise over n complex entries.
int gamma_field(qclhip_ctx* ctx, long long n, const double* SL, const double* SG, double* Gamma);
// Blocks for an elementwise grid-stride loop.
inline unsigned grid_blocks(long long n, int threads, int sm_count) {
    long long blocks = (n + threads - 1) / threads;
    long long cap = (long long)(sm_count > 0 ? sm_count : 32) * 32;
    if (blocks > cap) blocks = cap;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

// Grid of an (energy, column) kernel: x covers the NE energies with 256
// threads per block, y strides over the columns (about 16 blocks per CU).
inline dim3 ecol_grid(int NE, long long ncols, int sm_count) {
    const long long gx = ((long long)NE + 255) / 256;
    long long gy = (long long)(sm_count > 0 ? sm_count : 32) * 16 / (gx > 0 ? gx : 1);
    gy = gy < 1 ? 1 : (gy > ncols ? ncols : gy);
    gy = gy > 65535 ? 65535 : gy;
    return dim3((unsigned)(gx < 1 ? 1 : gx), (unsigned)(gy < 1 ? 1 : gy));
}

}  // namespace qclhip

#define QCLHIP_TRY_HIP(ctx, expr)                                              \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) return qclhip::hip_fail((ctx), _e, #expr);       \
    } while (0)
