// lb_probe.hip -- HBM bandwidth probe used to state the practical ceiling
// next to the nominal 8 TB/s (no counterpart in the reference).
//
// Kernels stream a buffer with 16-byte accesses, a grid-stride loop and
// enough work-groups to fill every CU: copy (read + write), read-only
// (reduced to one value per work-item so the loads are not dead),
// write-only, copy with non-temporal accesses, and the same with four loads
// in flight per work-item.  Each is timed with HIP events over `reps`
// launches.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ludwig_hip.h"

namespace {

__global__ void __launch_bounds__(256) probe_copy(const float4* __restrict__ a, float4* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) b[i] = a[i];
}

__global__ void __launch_bounds__(256) probe_read(const float4* __restrict__ a, float* __restrict__ out, int64_t n) {
    float s = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        float4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) probe_write(float4* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        b[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

// Copy with non-temporal loads and stores (the cache policy of the
// stream-collide kernel's outputs).
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) probe_copy_nt(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const v4f v = __builtin_nontemporal_load(a + i);
        __builtin_nontemporal_store(v, b + i);
    }
}

// Copy with four independent 16-byte loads in flight per work-item before
// the stores, non-temporal.  Elements past n are skipped by the bounds checks.
__global__ void __launch_bounds__(256) probe_copy_nt4(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += 4 * stride) {
        v4f v[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (i + u * stride < n) v[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (i + u * stride < n) __builtin_nontemporal_store(v[u], b + i + u * stride);
    }
}

}  // namespace

extern "C" int lb_bandwidth_probe(int device, int64_t bytes, int reps, double* gbps_out) {
    if (!gbps_out || bytes < (1 << 20) || reps < 1) return LB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return LB_EHIP;
    const int64_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    float* red = nullptr;
    int dev_cus = 0;
    if (hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return LB_EHIP;
    const unsigned groups = (unsigned)dev_cus * 8u;
    if (hipMalloc((void**)&a, (size_t)n * 16) != hipSuccess) return LB_ENOMEM;
    if (hipMalloc((void**)&b, (size_t)n * 16) != hipSuccess) {
        (void)hipFree(a);
        return LB_ENOMEM;
    }
    if (hipMalloc((void**)&red, (size_t)groups * 256 * 4) != hipSuccess) {
        (void)hipFree(a);
        (void)hipFree(b);
        return LB_ENOMEM;
    }
    int rc = LB_OK;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(probe_write, dim3(groups), dim3(256), 0, 0, a, n);
    for (int kind = 0; kind < 5 && rc == LB_OK; kind++) {
        // one untimed launch, then `reps` timed launches
        for (int r = -1; r < reps; r++) {
            if (r == 0) (void)hipEventRecord(e0, 0);
            if (kind == 0) hipLaunchKernelGGL(probe_copy, dim3(groups), dim3(256), 0, 0, a, b, n);
            if (kind == 1) hipLaunchKernelGGL(probe_read, dim3(groups), dim3(256), 0, 0, a, red, n);
            if (kind == 2) hipLaunchKernelGGL(probe_write, dim3(groups), dim3(256), 0, 0, b, n);
            if (kind == 3) hipLaunchKernelGGL(probe_copy_nt, dim3(groups), dim3(256), 0, 0, (const v4f*)a, (v4f*)b, n);
            if (kind == 4) hipLaunchKernelGGL(probe_copy_nt4, dim3(groups), dim3(256), 0, 0, (const v4f*)a, (v4f*)b, n);
        }
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) rc = LB_EHIP;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double moved = (double)n * 16.0 * (kind == 1 || kind == 2 ? 1.0 : 2.0) * reps;
        gbps_out[kind] = moved / (ms * 1e-3) / 1e9;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(red);
    return rc;
}
