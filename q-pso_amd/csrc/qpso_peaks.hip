// qpso_peaks -- the device's own ceilings for the render kernel's roofline,
// measured on the device it runs on (a measurement tool, not part of the
// library; built by `make -C q-pso_amd cal`).  Prints one JSON line:
//   hbm_copy_gbs      bytes read + written per second by a float4 copy of 1 GiB
//   lds_atomic_gops   32-bit LDS atomicMin operations per second, every lane of
//                     a wave on its own word (the raster's pixel writes)
//   lds_read_gbs      LDS bytes read per second with 128-bit reads
// Each kernel runs 5 timed launches after a warm-up; the best is reported.
// Every access is in bounds by construction: LDS indices are masked to the
// array, global indices are below the allocation's element count.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                               \
    do {                                                                                    \
        if ((x) != hipSuccess) {                                                            \
            std::fprintf(stderr, "qpso_peaks: %s failed at line %d\n", #x, __LINE__);       \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

namespace {

constexpr int kThreads = 1024;
constexpr int kWords = 8192;  // LDS words per workgroup (32 KiB)

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// 8 independent atomicMin per trip, lane l of a wave on word (l + 64 k) & mask:
// consecutive words, so the 64 lanes of one instruction meet no bank twice.
__global__ __launch_bounds__(kThreads) void k_lds_atomic(int* __restrict__ out, int trips) {
    __shared__ int lds[kWords];
    for (int i = threadIdx.x; i < kWords; i += kThreads) lds[i] = 0x7fffffff;
    __syncthreads();
    const int t = threadIdx.x;
    int v = 0x3fffffff - t;
    for (int r = 0; r < trips; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) atomicMin(&lds[(t + 1024 * k) & (kWords - 1)], v - k);
        v -= 8;
    }
    __syncthreads();
    if (t == 0) out[blockIdx.x] = lds[blockIdx.x & (kWords - 1)];
}

// 128-bit LDS reads, 8 per trip, summed so none is dead.
__global__ __launch_bounds__(kThreads) void k_lds_read(float* __restrict__ out, int trips) {
    __shared__ float4 lds[kWords / 4];
    for (int i = threadIdx.x; i < kWords / 4; i += kThreads) lds[i] = make_float4(1.0f, 2.0f, 3.0f, (float)i);
    __syncthreads();
    const int t = threadIdx.x;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int r = 0; r < trips; ++r) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float4 x = lds[(t + 256 * k + r) & (kWords / 4 - 1)];
            acc.x += x.x;
            acc.y += x.y;
            acc.z += x.z;
            acc.w += x.w;
        }
    }
    if (acc.x + acc.y + acc.z + acc.w == -1.0f) out[blockIdx.x] = acc.x;  // never true: keeps the reads live
}

template <class F>
int best_ms(F launch, hipEvent_t e0, hipEvent_t e1, float* ms) {
    launch();  // warm-up
    CK(hipDeviceSynchronize());
    *ms = 1e30f;
    for (int rep = 0; rep < 5; ++rep) {
        CK(hipEventRecord(e0, 0));
        launch();
        CK(hipEventRecord(e1, 0));
        CK(hipEventSynchronize(e1));
        float m = 0.0f;
        CK(hipEventElapsedTime(&m, e0, e1));
        if (m < *ms) *ms = m;
    }
    return 0;
}

}  // namespace

int main() {
    int n_cu = 0;
    CK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, 0));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));

    const size_t bytes = 1ull << 30, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 0, bytes));
    float ms_copy = 0.0f;
    if (best_ms([&] { hipLaunchKernelGGL(k_copy, dim3(n_cu * 64), dim3(256), 0, 0, a, b, n); }, e0, e1, &ms_copy))
        return 1;
    CK(hipFree(a));
    CK(hipFree(b));

    const int blocks = n_cu * 4, trips = 4096;
    int* iout = nullptr;
    float* fout = nullptr;
    CK(hipMalloc(&iout, sizeof(int) * blocks));
    CK(hipMalloc(&fout, sizeof(float) * blocks));
    float ms_atomic = 0.0f, ms_read = 0.0f;
    if (best_ms([&] { hipLaunchKernelGGL(k_lds_atomic, dim3(blocks), dim3(kThreads), 0, 0, iout, trips); }, e0, e1,
                &ms_atomic))
        return 1;
    if (best_ms([&] { hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(kThreads), 0, 0, fout, trips); }, e0, e1,
                &ms_read))
        return 1;
    CK(hipFree(iout));
    CK(hipFree(fout));

    const double ops = (double)blocks * kThreads * trips * 8;
    std::printf("{\"n_cu\": %d, \"hbm_copy_gbs\": %.1f, \"lds_atomic_gops\": %.1f, \"lds_read_gbs\": %.1f}\n", n_cu,
                2.0 * (double)bytes / (ms_copy * 1e-3) / 1e9, ops / (ms_atomic * 1e-3) / 1e9,
                ops * 16.0 / (ms_read * 1e-3) / 1e9);
    return 0;
}
