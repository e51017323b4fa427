// hbm_copy_probe.hip -- device-to-device copy rate of this GPU's HBM, the ceiling
// that the step kernels' traffic is compared against.  Copies 1 GiB with 4-,
// 8- and 16-byte accesses per lane, default and nontemporal cache policy,
// either one element per thread (grid = elements / 256) or a grid-stride loop
// over a fixed number of workgroups; prints read+write bytes per second.
// Then the step's own access shape without its arithmetic: 256^3 nodes, ten
// FP32 values read and written per node (nontemporal), 32x16-lane workgroups
// marching 32 planes with the next plane prefetched, for three layouts of the
// ten values: ten arrays of 4 B (the solver's), five arrays of interleaved
// pairs (8 B per lane), two 16 B groups and one 8 B group.
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

constexpr int SX = 256, SY = 256, SZ = 256, STX = 32, STY = 16, SLZ = 32;
constexpr size_t SN = (size_t)SX * SY * SZ;

// node index of plane z for this lane: tiled order [z][tile][STY][STX]
__device__ __forceinline__ unsigned stream_node(int z) {
    constexpr int nt = (SX / STX) * (SY / STY);
    return (unsigned)((z * nt + (int)(blockIdx.x % nt)) * (STX * STY) + (int)threadIdx.x);
}

// L = 0: ten float arrays; 1: five arrays of float pairs; 2: 16 B + 16 B + 8 B groups
template <int L>
__global__ __launch_bounds__(512) void stream_kernel(const char *__restrict__ in, char *__restrict__ out) {
    const int z0 = (int)(blockIdx.x / ((SX / STX) * (SY / STY))) * SLZ;
    float cur[10], nxt[10];
    auto load = [&](int z, float (&m)[10]) {
        const unsigned i = stream_node(z);
        if constexpr (L == 0) {
            for (int k = 0; k < 10; k++) m[k] = __builtin_nontemporal_load((const float *)(in + (size_t)k * SN * 4) + i);
        } else if constexpr (L == 1) {
            for (int g = 0; g < 5; g++) {
                const v2f v = __builtin_nontemporal_load((const v2f *)(in + (size_t)g * SN * 8) + i);
                m[2 * g] = v.x;
                m[2 * g + 1] = v.y;
            }
        } else {
            const v4f a = __builtin_nontemporal_load((const v4f *)in + i);
            const v4f b = __builtin_nontemporal_load((const v4f *)(in + SN * 16) + i);
            const v2f d = __builtin_nontemporal_load((const v2f *)(in + SN * 32) + i);
            m[0] = a.x, m[1] = a.y, m[2] = a.z, m[3] = a.w, m[4] = b.x, m[5] = b.y, m[6] = b.z, m[7] = b.w;
            m[8] = d.x, m[9] = d.y;
        }
    };
    auto store = [&](int z, const float (&m)[10]) {
        const unsigned i = stream_node(z);
        if constexpr (L == 0) {
            for (int k = 0; k < 10; k++) __builtin_nontemporal_store(m[k], (float *)(out + (size_t)k * SN * 4) + i);
        } else if constexpr (L == 1) {
            for (int g = 0; g < 5; g++)
                __builtin_nontemporal_store(v2f{m[2 * g], m[2 * g + 1]}, (v2f *)(out + (size_t)g * SN * 8) + i);
        } else {
            __builtin_nontemporal_store(v4f{m[0], m[1], m[2], m[3]}, (v4f *)out + i);
            __builtin_nontemporal_store(v4f{m[4], m[5], m[6], m[7]}, (v4f *)(out + SN * 16) + i);
            __builtin_nontemporal_store(v2f{m[8], m[9]}, (v2f *)(out + SN * 32) + i);
        }
    };
    load(z0, cur);
    for (int z = z0; z < z0 + SLZ; z++) {
        if (z + 1 < z0 + SLZ) load(z + 1, nxt);
        store(z, cur);
        for (int k = 0; k < 10; k++) cur[k] = nxt[k];
    }
}

// microseconds per launch (20 launches after 3 warm-up launches)
template <int L>
double stream_us(const void *in, void *out) {
    const unsigned blocks = (SX / STX) * (SY / STY) * (SZ / SLZ);
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    for (int w = 0; w < 3; w++) stream_kernel<L><<<blocks, 512>>>((const char *)in, (char *)out);
    CK(hipEventRecord(e0));
    for (int r = 0; r < 20; r++) stream_kernel<L><<<blocks, 512>>>((const char *)in, (char *)out);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipGetLastError());
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return ms * 1e3 / 20;
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

    void *in = nullptr, *out = nullptr;
    CK(hipMalloc(&in, 40 * SN));
    CK(hipMalloc(&out, 40 * SN));
    CK(hipMemset(in, 0, 40 * SN));
    CK(hipMemset(out, 0, 40 * SN));
    static const char *names[3] = {"ten 4 B arrays", "five 8 B pair arrays", "16 B + 16 B + 8 B groups"};
    std::printf("step-shaped stream, 256^3 x 10 FP32 read + written (80 B/node), 1024 workgroups x 32 planes\n");
    for (int rep = 0; rep < 2; rep++) {
        const double us[3] = {stream_us<0>(in, out), stream_us<1>(in, out), stream_us<2>(in, out)};
        for (int l = 0; l < 3; l++)
            std::printf("  %-26s %7.1f us  %.3f TB/s\n", names[l], us[l], 80.0 * SN / (us[l] * 1e-6) / 1e12);
        std::fflush(stdout);
    }
    CK(hipFree(in));
    CK(hipFree(out));
    return 0;
}
