// hbm_copy_probe.hip -- device-to-device copy rate of this GPU's HBM, the ceiling
// that the step kernels' traffic is compared against.  Copies 1 GiB with 4-,
// 8- and 16-byte accesses per lane, default and nontemporal cache policy,
// either one element per thread (grid = elements / 256) or a grid-stride loop
// over a fixed number of workgroups; prints read+write bytes per second.
//   cudalbm_amd/bin/hbmCopyProbe
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

void check(hipError_t e, int line) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error %s (line %d)\n", hipGetErrorString(e), line);
        std::exit(1);
    }
}
#define CK(x) check((x), __LINE__)

template <typename T, bool NT>
__global__ void copy_kernel(const T *__restrict__ a, T *__restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if constexpr (NT)
            __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
        else
            b[i] = a[i];
    }
}

// TB/s of read + write over `reps` launches after 3 warm-up launches;
// blocks = 0: one element per thread
template <typename T, bool NT>
double rate(const void *a, void *b, size_t bytes, size_t blocks) {
    const size_t n = bytes / sizeof(T);
    if (blocks == 0) blocks = n / 256;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    for (int w = 0; w < 3; w++) copy_kernel<T, NT><<<(unsigned)blocks, 256>>>((const T *)a, (T *)b, n);
    const int reps = 20;
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) copy_kernel<T, NT><<<(unsigned)blocks, 256>>>((const T *)a, (T *)b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipGetLastError());
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e12;
}

}  // namespace

int main() {
    const size_t bytes = (size_t)1 << 30;
    void *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMemset(a, 0, bytes));
    CK(hipMemset(b, 0, bytes));
    std::printf("1 GiB device copy, TB/s of read + write (4/8/16 B per lane, default | nontemporal policy)\n");
    for (size_t blocks : {(size_t)0, (size_t)131072, (size_t)32768, (size_t)8192}) {
        if (blocks == 0)
            std::printf("one element per thread    ");
        else
            std::printf("grid-stride, %6zu groups ", blocks);
        std::printf(" 4 B %.3f | %.3f   8 B %.3f | %.3f   16 B %.3f | %.3f\n", rate<float, false>(a, b, bytes, blocks),
                    rate<float, true>(a, b, bytes, blocks), rate<v2f, false>(a, b, bytes, blocks),
                    rate<v2f, true>(a, b, bytes, blocks), rate<v4f, false>(a, b, bytes, blocks),
                    rate<v4f, true>(a, b, bytes, blocks));
        std::fflush(stdout);
    }
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
