// lb_probe.hip -- HBM bandwidth probe used to state the practical ceiling
// next to the nominal 8 TB/s (no counterpart in the reference).
//
// Plain streaming kernels over two buffers, 16-byte accesses, one element per
// work-item and a grid sized to the buffer (no grid-stride loop): on MI355X
// this is the fastest form measured -- consecutive work-groups cover
// consecutive 4 KiB of memory.  Sweeps over 1-8 elements per work-item and
// grid-stride grids of 4-64 work-groups per CU were all slower (copy 4.3-6.1
// TB/s against 6.6 TB/s).  Kernels: copy (read + write), read-only (a value
// is written only if the sum hits an impossible bit pattern, so the loads are
// live), write-only, copy with non-temporal loads and stores, and the
// grid-stride copy (8 work-groups per CU) for comparison.  Each is timed with
// HIP events over `reps` launches after one untimed launch.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ludwig_hip.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) probe_copy(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void __launch_bounds__(256) probe_read(const v4f* __restrict__ a, float* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const v4f v = __builtin_nontemporal_load(a + i);
        const float s = v.x + v.y + v.z + v.w;
        if (__float_as_uint(s) == 0x7f800001u) out[threadIdx.x] = s;  // a NaN pattern the data never holds
    }
}

__global__ void __launch_bounds__(256) probe_write(v4f* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const v4f v = {1.0f, 2.0f, 3.0f, 4.0f};
    if (i < n) __builtin_nontemporal_store(v, b + i);
}

// Copy with non-temporal loads and stores (the cache policy of the
// stream-collide kernel's outputs).
__global__ void __launch_bounds__(256) probe_copy_nt(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

// Grid-stride copy, non-temporal (8 work-groups per CU).
__global__ void __launch_bounds__(256) probe_copy_stride(const v4f* __restrict__ a, v4f* __restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

}  // namespace

extern "C" int lb_bandwidth_probe(int device, int64_t bytes, int reps, double* gbps_out) {
    if (!gbps_out || bytes < (1 << 20) || reps < 1) return LB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return LB_EHIP;
    const int64_t n = bytes / 16;
    v4f *a = nullptr, *b = nullptr;
    float* red = nullptr;
    int dev_cus = 0;
    if (hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return LB_EHIP;
    const unsigned sized = (unsigned)((n + 255) / 256);
    const unsigned strided = (unsigned)dev_cus * 8u;
    if (hipMalloc((void**)&a, (size_t)n * 16) != hipSuccess) return LB_ENOMEM;
    if (hipMalloc((void**)&b, (size_t)n * 16) != hipSuccess) {
        (void)hipFree(a);
        return LB_ENOMEM;
    }
    if (hipMalloc((void**)&red, 256 * sizeof(float)) != hipSuccess) {
        (void)hipFree(a);
        (void)hipFree(b);
        return LB_ENOMEM;
    }
    int rc = LB_OK;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(probe_write, dim3(sized), dim3(256), 0, 0, a, n);
    for (int kind = 0; kind < 5 && rc == LB_OK; kind++) {
        // one untimed launch, then `reps` timed launches
        for (int r = -1; r < reps; r++) {
            if (r == 0) (void)hipEventRecord(e0, 0);
            if (kind == 0) hipLaunchKernelGGL(probe_copy, dim3(sized), dim3(256), 0, 0, a, b, n);
            if (kind == 1) hipLaunchKernelGGL(probe_read, dim3(sized), dim3(256), 0, 0, a, red, n);
            if (kind == 2) hipLaunchKernelGGL(probe_write, dim3(sized), dim3(256), 0, 0, b, n);
            if (kind == 3) hipLaunchKernelGGL(probe_copy_nt, dim3(sized), dim3(256), 0, 0, a, b, n);
            if (kind == 4) hipLaunchKernelGGL(probe_copy_stride, dim3(strided), dim3(256), 0, 0, a, b, n);
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
