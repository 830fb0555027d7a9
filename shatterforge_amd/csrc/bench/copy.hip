// Bench-only program, built apart from the library (shatterforge_amd/sfb_copy): a device-to-device copy of known size,
// the reference reading for the HBM byte counters (FETCH_SIZE and WRITE_SIZE per byte moved), profiled in a run of its
// own (rocprofv3 --pmc FETCH_SIZE -- shatterforge_amd/sfb_copy). Nothing in the library calls it.
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {
__global__ void __launch_bounds__(256) k_fill16(uint4* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}
// one 16-byte load and one 16-byte store per element, grid-stride
__global__ void __launch_bounds__(256) k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

// Copies `bytes` (a multiple of 16) `reps` times between two fresh device buffers; *ms: the copies' time by HIP
// events. Returns 0, or -1 when a HIP call fails.
static int sfb_copy(size_t bytes, int reps, double* ms) {
    if (bytes == 0 || bytes % 16 || reps < 1 || !ms) return -1;
    const size_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -1;
    float t = 0.0f;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    hipLaunchKernelGGL(k_fill16, dim3(4096), dim3(256), 0, 0, a, n);
    if (hipDeviceSynchronize() != hipSuccess) goto done;
    if (hipEventRecord(e0, 0) != hipSuccess) goto done;
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy16, dim3(8192), dim3(256), 0, 0, b, a, n);
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) goto done;
    *ms = t;
    rc = 0;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

// sfb_copy [GiB] [copies]: prints the bytes per copy and the read + write rate
int main(int argc, char** argv) {
    const double gib = argc > 1 ? atof(argv[1]) : 1.0;
    const int reps = argc > 2 ? atoi(argv[2]) : 5;
    const size_t bytes = (size_t)(gib * (double)(1ull << 30)) / 16 * 16;
    double ms = 0.0;
    if (sfb_copy(bytes, reps, &ms) != 0) { fprintf(stderr, "copy failed\n"); return 1; }
    printf("{\"kernel\": \"k_copy16\", \"bytes_per_copy\": %zu, \"copies\": %d, \"GBs_read_plus_write\": %.1f}\n", bytes, reps,
           2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9);
    return 0;
}
