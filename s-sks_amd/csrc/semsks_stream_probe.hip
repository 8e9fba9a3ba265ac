// semsks_stream_probe: practical HBM ceilings for the step kernel's access mix.
//
// The step kernel reads u(t) and u(t-dt) (12 B per node each) and the packed
// compensation word of level t (4 B), and writes u(t+dt) (12 B) and its word
// (4 B): 28 B read and 16 B written per node, five streams.  This program
// times, with HIP events, grid-stride kernels that move exactly that mix
// ("step_mix"), a device copy and a triad a = b + 2c over the same memory,
// each the best of 6 launches at 2, 4 and 8 blocks of 256 threads per CU.
// It prints one JSON object.
//
//   semsks_stream_probe [nodes]      (default 2^26 nodes: 2.95 GB of planes)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                           \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_));    \
            exit(1);                                                                     \
        }                                                                                \
    } while (0)

/* one 4-node group per iteration: three float4 of each xyz plane, one uint4 of words */
__global__ __launch_bounds__(256) void step_mix(const float4 *__restrict__ u0, const float4 *__restrict__ u1,
                                                const uint4 *__restrict__ w0, float4 *__restrict__ u2,
                                                uint4 *__restrict__ w2, size_t ngrp)
{
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < ngrp; g += (size_t)gridDim.x * blockDim.x) {
        const uint4 w = w0[g];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float4 x = u0[3 * g + j], y = u1[3 * g + j];
            u2[3 * g + j] = make_float4(2.f * x.x - y.x, 2.f * x.y - y.y, 2.f * x.z - y.z, 2.f * x.w - y.w);
        }
        w2[g] = make_uint4(w.x + 1u, w.y + 1u, w.z + 1u, w.w + 1u);
    }
}

__global__ __launch_bounds__(256) void copy4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ __launch_bounds__(256) void triad4(float4 *__restrict__ a, const float4 *__restrict__ b,
                                              const float4 *__restrict__ c, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 x = b[i], y = c[i];
        a[i] = make_float4(x.x + 2.f * y.x, x.y + 2.f * y.y, x.z + 2.f * y.z, x.w + 2.f * y.w);
    }
}

template <class F>
static void best_of(const char *name, double bytes, int ncu, F launch, bool last)
{
    hipEvent_t a, b;
    CHK(hipEventCreate(&a));
    CHK(hipEventCreate(&b));
    float best = 1e30f;
    int best_bpc = 0;
    for (int bpc : {2, 4, 8}) {
        for (int r = 0; r < 6; ++r) {
            CHK(hipEventRecord(a));
            launch(ncu * bpc);
            CHK(hipGetLastError());
            CHK(hipEventRecord(b));
            CHK(hipEventSynchronize(b));
            float ms = 0.f;
            CHK(hipEventElapsedTime(&ms, a, b));
            if (r >= 1 && ms < best) { best = ms; best_bpc = bpc; }
        }
    }
    printf("  \"%s\": {\"bytes\": %.0f, \"best_ms\": %.4f, \"GB_per_s\": %.1f, \"blocks_per_cu\": %d}%s\n", name, bytes,
           best, bytes / (best * 1e-3) / 1e9, best_bpc, last ? "" : ",");
    CHK(hipEventDestroy(a));
    CHK(hipEventDestroy(b));
}

int main(int argc, char **argv)
{
    const size_t nodes = argc > 1 ? (size_t)atoll(argv[1]) / 64 * 64 : ((size_t)1 << 26);
    if (nodes < 64) {
        fprintf(stderr, "need at least 64 nodes\n");
        return 2;
    }
    const size_t ngrp = nodes / 4;
    hipDeviceProp_t prop;
    CHK(hipGetDeviceProperties(&prop, 0));
    const int ncu = prop.multiProcessorCount;
    /* three xyz planes and two word planes, each 256-byte aligned by hipMalloc */
    float4 *u0, *u1, *u2;
    uint4 *w0, *w2;
    CHK(hipMalloc(&u0, nodes * 12));
    CHK(hipMalloc(&u1, nodes * 12));
    CHK(hipMalloc(&u2, nodes * 12));
    CHK(hipMalloc(&w0, nodes * 4));
    CHK(hipMalloc(&w2, nodes * 4));
    CHK(hipMemset(u0, 0, nodes * 12));
    CHK(hipMemset(u1, 0, nodes * 12));
    CHK(hipMemset(u2, 0, nodes * 12));
    CHK(hipMemset(w0, 0, nodes * 4));
    CHK(hipMemset(w2, 0, nodes * 4));
    CHK(hipDeviceSynchronize());
    const size_t n4 = nodes * 3 / 4;   /* float4 per xyz plane */
    printf("{\n  \"nodes\": %zu, \"cus\": %d,\n", nodes, ncu);
    best_of("step_mix", 44.0 * nodes, ncu, [&](int nb) {
        hipLaunchKernelGGL(step_mix, dim3(nb), dim3(256), 0, 0, u0, u1, w0, u2, w2, ngrp);
    }, false);
    best_of("copy", 32.0 * n4, ncu, [&](int nb) { hipLaunchKernelGGL(copy4, dim3(nb), dim3(256), 0, 0, u0, u1, n4); },
            false);
    best_of("triad", 48.0 * n4, ncu, [&](int nb) {
        hipLaunchKernelGGL(triad4, dim3(nb), dim3(256), 0, 0, u2, u0, u1, n4);
    }, true);
    printf("}\n");
    CHK(hipFree(u0));
    CHK(hipFree(u1));
    CHK(hipFree(u2));
    CHK(hipFree(w0));
    CHK(hipFree(w2));
    return 0;
}
