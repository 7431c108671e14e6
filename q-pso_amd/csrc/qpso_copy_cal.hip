// qpso_copy_cal -- measuring stick for rocprofv3's FETCH_SIZE and WRITE_SIZE
// counters: a grid-stride float4 copy of a known 256 MiB, launched 5 times.
// The ratio of the bytes copied to what the counters report for one launch
// scales the counters read on the render+fitness kernel to HBM bytes.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main() {
    const size_t bytes = 256ull << 20, n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(copy4, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 2;
    std::printf("qpso_copy_cal: 5 x %zu bytes copied\n", bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
