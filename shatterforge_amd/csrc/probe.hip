// Measurement probe: a plain device-to-device copy of a known number of bytes. Its HBM counter readings (FETCH_SIZE,
// WRITE_SIZE) per byte moved calibrate the counters that measure the hot kernels' traffic.
#include "sf_host.h"
#include "../../include/sf.h"
#include <hip/hip_runtime.h>

namespace {
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}
}  // namespace

extern "C" int sf_probe_copy(int64_t bytes, int reps, float* ms_out) {
    if (bytes < 16 || reps < 1) { sf::set_error(SF_E_ARG, "bad size"); return SF_E_ARG; }
    size_t n = (size_t)bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    SF_HIP(hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) { (void)hipFree(a); sf::set_error(SF_E_HIP, "hipMalloc"); return SF_E_HIP; }
    SF_HIP(hipMemset(a, 0, n * 16));
    SF_HIP(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    SF_HIP(hipEventCreate(&e0));
    SF_HIP(hipEventCreate(&e1));
    SF_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(256), 0, 0, a, b, n);
    SF_HIP(hipEventRecord(e1, 0));
    SF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    SF_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_out) *ms_out = ms;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    (void)hipFree(a); (void)hipFree(b);
    return 0;
}
