// gslam_stream_probe: the streaming bandwidth this GPU reaches for working
// sets of a given size, re-read by every launch as the PCG iterations re-read
// their matrix and vectors.  For each size (MB, decimal): a read kernel (16
// bytes per lane per step, grid-stride, one partial per workgroup) and a copy
// kernel, each timed with HIP events over back-to-back launches.  It sets the
// ceiling that the PCG kernels' achieved bytes per second are read against:
// a working set that spans the eight 4 MB L2s reads faster than HBM, one of
// 69 MB (the one-launch iteration at N = 100k) does not.
//
//   gslam_stream_probe [MB ...]          (default: 31 69 1024)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define PROBE_CHECK(x)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "gslam_stream_probe: %s (line %d)\n", hipGetErrorString(e_), \
                         __LINE__);                                                      \
            std::exit(1);                                                                \
        }                                                                                \
    } while (0)

constexpr int kBlock = 256;
constexpr int kGrid = 2048;  // 8 workgroups per CU

__global__ void __launch_bounds__(kBlock)
stream_read_kernel(const double2 *__restrict__ a, int64_t n2, double *__restrict__ part)
{
    double s = 0.0;
    for (int64_t i = blockIdx.x * int64_t{kBlock} + threadIdx.x; i < n2;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const double2 v = a[i];
        s += v.x + v.y;
    }
    __shared__ double red[kBlock];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = kBlock / 2; k > 0; k >>= 1) {
        if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kBlock)
stream_copy_kernel(const double2 *__restrict__ a, double2 *__restrict__ b, int64_t n2)
{
    for (int64_t i = blockIdx.x * int64_t{kBlock} + threadIdx.x; i < n2;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
        b[i] = a[i];
}

int main(int argc, char **argv)
{
    std::vector<long> sizes;
    for (int i = 1; i < argc; ++i) sizes.push_back(std::atol(argv[i]));
    if (sizes.empty()) sizes = {31, 69, 1024};
    double *part = nullptr;
    PROBE_CHECK(hipMalloc(&part, kGrid * sizeof(double)));
    hipEvent_t e0, e1;
    PROBE_CHECK(hipEventCreate(&e0));
    PROBE_CHECK(hipEventCreate(&e1));
    for (long mb : sizes) {
        if (mb < 1 || mb > 16384) {
            std::fprintf(stderr, "gslam_stream_probe: size %ld MB out of range [1, 16384]\n", mb);
            return 2;
        }
        const int64_t n2 = mb * int64_t{1000000} / 16;
        const double bytes = 16.0 * static_cast<double>(n2);
        double2 *a = nullptr, *b = nullptr;
        PROBE_CHECK(hipMalloc(&a, n2 * 16));
        PROBE_CHECK(hipMalloc(&b, n2 * 16));
        PROBE_CHECK(hipMemset(a, 0, n2 * 16));
        const int reps = mb > 256 ? 20 : 200;
        float ms = 0.f;
        for (int w = 0; w < 5; ++w)
            hipLaunchKernelGGL(stream_read_kernel, dim3(kGrid), dim3(kBlock), 0, 0, a, n2, part);
        PROBE_CHECK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(stream_read_kernel, dim3(kGrid), dim3(kBlock), 0, 0, a, n2, part);
        PROBE_CHECK(hipEventRecord(e1, 0));
        PROBE_CHECK(hipEventSynchronize(e1));
        PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double us_read = 1000.0 * ms / reps;
        for (int w = 0; w < 5; ++w)
            hipLaunchKernelGGL(stream_copy_kernel, dim3(kGrid), dim3(kBlock), 0, 0, a, b, n2);
        PROBE_CHECK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(stream_copy_kernel, dim3(kGrid), dim3(kBlock), 0, 0, a, b, n2);
        PROBE_CHECK(hipEventRecord(e1, 0));
        PROBE_CHECK(hipEventSynchronize(e1));
        PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double us_copy = 1000.0 * ms / reps;
        PROBE_CHECK(hipGetLastError());
        std::printf("stream %ld MB: read %.2f us %.2f TB/s; copy %.2f us %.2f TB/s read+write\n", mb,
                    us_read, bytes / us_read / 1e6, us_copy, 2.0 * bytes / us_copy / 1e6);
        std::fflush(stdout);
        PROBE_CHECK(hipFree(a));
        PROBE_CHECK(hipFree(b));
    }
    PROBE_CHECK(hipFree(part));
    return 0;
}
