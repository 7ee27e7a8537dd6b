// Streaming kernels that measure the achievable HBM rates of the device the
// apply runs on: a read-only sweep, a copy and a write-only sweep over
// buffers far larger than the caches.  They are the stated ceilings of the
// apply's roofline (its traffic is mostly reads plus one write of each
// output) and the calibration of the PMC byte counters (a copy of known
// size).  Measurement only: built into libcrhbm.so, not the regridding library.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kThreads = 256;
constexpr int kUnroll = 4;

// Grid-stride loop, kUnroll independent 16-byte loads in flight per lane.
// The sum is stored only if it equals a value the callers' inputs never give
// (sums of small integers), so the loads stay and nothing is written.
__global__ void __launch_bounds__(kThreads) k_read(const v4f *__restrict__ a, int64_t n4, float *__restrict__ sink) {
    const int64_t stride = (int64_t)gridDim.x * kThreads * kUnroll;
    float acc = 0.f;
    for (int64_t base = (int64_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; base < n4; base += stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t i = base + (int64_t)u * kThreads;
            v[u] = i < n4 ? a[i] : v4f{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) acc += v[u].x + v[u].y + v[u].z + v[u].w;
    }
    if (acc == 0.375f) sink[threadIdx.x] = acc;
}

__global__ void __launch_bounds__(kThreads) k_copy(const v4f *__restrict__ a, v4f *__restrict__ b, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * kThreads * kUnroll;
    for (int64_t base = (int64_t)blockIdx.x * kThreads * kUnroll + threadIdx.x; base < n4; base += stride) {
        v4f v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t i = base + (int64_t)u * kThreads;
            if (i < n4) v[u] = a[i];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int64_t i = base + (int64_t)u * kThreads;
            if (i < n4) __builtin_nontemporal_store(v[u], b + i);
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_write(v4f *__restrict__ b, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    const v4f z = v4f{1.f, 2.f, 3.f, 4.f};
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride)
        __builtin_nontemporal_store(z, b + i);
}

}  // namespace

extern "C" {

// kind 0: read n bytes of a; 1: copy n bytes a -> b; 2: write n bytes of b.
// n must be a multiple of 16 and the buffers 16-byte aligned; sink holds 256
// floats.  blocks: workgroups of 256 threads (grid-stride).  Returns a hipError_t.
int crhbm_launch(int kind, const void *a, void *b, int64_t nbytes, void *sink, int blocks, void *stream) {
    if (nbytes <= 0 || nbytes % 16 != 0 || blocks <= 0) return (int)hipErrorInvalidValue;
    const int64_t n4 = nbytes / 16;
    hipStream_t s = (hipStream_t)stream;
    if (kind == 0) {
        if (!a || !sink) return (int)hipErrorInvalidValue;
        hipLaunchKernelGGL(k_read, dim3(blocks), dim3(kThreads), 0, s, (const v4f *)a, n4, (float *)sink);
    } else if (kind == 1) {
        if (!a || !b) return (int)hipErrorInvalidValue;
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kThreads), 0, s, (const v4f *)a, (v4f *)b, n4);
    } else if (kind == 2) {
        if (!b) return (int)hipErrorInvalidValue;
        hipLaunchKernelGGL(k_write, dim3(blocks), dim3(kThreads), 0, s, (v4f *)b, n4);
    } else {
        return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // extern "C"
