// hbm_copy: a plain device-to-device copy of a known number of bytes.
//
// Used to calibrate the HBM traffic counters (FETCH_SIZE / WRITE_SIZE of
// rocprofv3 --pmc read per dispatch for exactly `bytes` read and `bytes`
// written) and to measure the practical copy bandwidth of the device, the
// ceiling the memory-bound passes are compared with.
//
//   hbm_copy [MiB=1024] [reps=20]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

#define CHK(x)                                                                    \
    do {                                                                          \
        hipError_t e = (x);                                                       \
        if (e != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));           \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main(int argc, char **argv) {
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const long bytes = mib << 20;
    const long n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 1, bytes));
    int dev = 0, cus = 0;
    CHK(hipGetDevice(&dev));
    CHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = cus * 8;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e1, 0));
    CHK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / reps;
    std::printf("hbm_copy: %ld bytes read + %ld bytes written per dispatch, %.4f ms, %.1f GB/s (read+write)\n", bytes,
                bytes, per, 2.0 * bytes / (per * 1e-3) / 1e9);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
