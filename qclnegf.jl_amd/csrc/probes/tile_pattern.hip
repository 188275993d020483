// HBM rate of the per-block tile access pattern of the GF-solve kernels,
// against plain streaming: 5 fields read and 2 written (the Keldysh step's
// byte mix), N = 24, NE = 2048, Nk = 64, no LDS and no arithmetic beyond the
// sums. A block's element (i,j) of energy e sits at e + NE Nk (i + N j), so
// a tile of W consecutive energies moves N^2 segments of 16W bytes, 2 MB
// apart (a power of two at this size; the padded runs put a gap behind every
// plane). Standalone timing program (not part of the library):
//   hipcc --offload-arch=gfx950 -O3 -o tile_pattern tile_pattern.hip && ./tile_pattern
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

struct c64 {
    double x, y;
};
__device__ inline c64 add(c64 a, c64 b) { return {a.x + b.x, a.y + b.y}; }

__global__ void k_stream(long long n, const c64* a, const c64* b, const c64* c, const c64* d, const c64* e, c64* o0,
                         c64* o1) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        o0[i] = add(add(a[i], b[i]), c[i]);
        o1[i] = add(d[i], e[i]);
    }
}

// Consecutive tiles on one XCD (blocks are dispatched round-robin over 8).
__device__ inline long long xcd_order(long long b, long long nb) {
    const long long q = nb >> 3, rem = nb & 7, x = b & 7, s = b >> 3;
    return x * q + (x < rem ? x : rem) + s;
}

// One workgroup of 64*WAVES threads per tile of W consecutive blocks
// (W a multiple of WAVES); thread f moves energy f % W of element f / W.
// plane: distance between the planes (i,j) in elements (nblk, or nblk plus
// a pad that breaks the power-of-two stride of NE = 2048, Nk = 64).
template <int W, int WAVES, bool XCD>
__global__ __launch_bounds__(64 * WAVES) void k_tile(long long nblk, int N, const c64* a, const c64* b, const c64* c,
                                                     const c64* d, const c64* e, c64* o0, c64* o1, long long plane) {
    const long long tile = XCD ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x;
    const long long blk0 = tile * W;
    const int total = N * N * W;
    for (int f = threadIdx.x; f < total; f += 64 * WAVES) {
        const int w = f % W, ij = f / W;
        const size_t idx = blk0 + w + (size_t)plane * ij;
        o0[idx] = add(add(a[idx], b[idx]), c[idx]);
        o1[idx] = add(d[idx], e[idx]);
    }
}

#define CK(x)                                                                              \
    do {                                                                                   \
        hipError_t r_ = (x);                                                               \
        if (r_ != hipSuccess) {                                                            \
            printf("HIP error %s at line %d\n", hipGetErrorString(r_), __LINE__);           \
            exit(1);                                                                       \
        }                                                                                  \
    } while (0)

template <typename F>
double time_ms(F launch) {
    hipEvent_t s, t;
    CK(hipEventCreate(&s));
    CK(hipEventCreate(&t));
    launch();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(s));
    for (int r = 0; r < 10; ++r) launch();
    CK(hipEventRecord(t));
    CK(hipEventSynchronize(t));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, s, t));
    return ms / 10;
}

int main() {
    const int N = 24, NE = 2048, Nk = 64;
    const long long nblk = (long long)NE * Nk, n = nblk * N * N;
    const long long pad = 256;  // elements (4 KiB) appended to every plane in the padded runs
    const size_t bytes = (size_t)(nblk + pad) * N * N * sizeof(c64);
    c64* f[7];
    for (int i = 0; i < 7; ++i) {
        CK(hipMalloc(&f[i], bytes));
        CK(hipMemset(f[i], 0, bytes));
    }
    const double traffic = 7.0 * n * sizeof(c64);
    printf("N=%d NE=%d Nk=%d: field %.4e B, 5 read + 2 written = %.4e B per pass (mean of 10)\n", N, NE, Nk,
           (double)(n * sizeof(c64)), traffic);
    auto report = [&](const char* name, double ms) { printf("%-34s %7.3f ms  %5.2f TB/s\n", name, ms, traffic / ms / 1e9); };
#define ARGS nblk, N, f[0], f[1], f[2], f[3], f[4], f[5], f[6], nblk
#define ARGS_PAD(p) nblk, N, f[0], f[1], f[2], f[3], f[4], f[5], f[6], nblk + (p)
    report("streaming", time_ms([&] { k_stream<<<2048 * 8, 256>>>(n, f[0], f[1], f[2], f[3], f[4], f[5], f[6]); }));
    report("tile W=1  (16 B segments)", time_ms([&] { k_tile<1, 1, false><<<nblk, 64>>>(ARGS); }));
    report("tile W=2  (32 B)", time_ms([&] { k_tile<2, 2, false><<<nblk / 2, 128>>>(ARGS); }));
    report("tile W=4  (64 B)", time_ms([&] { k_tile<4, 4, false><<<nblk / 4, 256>>>(ARGS); }));
    report("tile W=4  (64 B), XCD-aware order", time_ms([&] { k_tile<4, 4, true><<<nblk / 4, 256>>>(ARGS); }));
    report("tile W=8  (128 B)", time_ms([&] { k_tile<8, 8, false><<<nblk / 8, 512>>>(ARGS); }));
    report("tile W=8  (128 B) by 4 waves", time_ms([&] { k_tile<8, 4, false><<<nblk / 8, 256>>>(ARGS); }));
    report("tile W=16 (256 B)", time_ms([&] { k_tile<16, 16, false><<<nblk / 16, 1024>>>(ARGS); }));
    report("tile W=4, XCD order, planes +256 B", time_ms([&] { k_tile<4, 4, true><<<nblk / 4, 256>>>(ARGS_PAD(16)); }));
    report("tile W=4, XCD order, planes +4 KiB", time_ms([&] { k_tile<4, 4, true><<<nblk / 4, 256>>>(ARGS_PAD(pad)); }));
    report("tile W=8, planes +4 KiB", time_ms([&] { k_tile<8, 8, false><<<nblk / 8, 512>>>(ARGS_PAD(pad)); }));
#undef ARGS
#undef ARGS_PAD
    for (int i = 0; i < 7; ++i) CK(hipFree(f[i]));
    return 0;
}
