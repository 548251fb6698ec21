// Streaming probe kernels for HBM counter calibration and the measured-bandwidth line of the
// benchmark (libdugks_probe.so; not part of the DUGKS C-ABI).
//
// k_probe_read<W> reads a buffer once with coalesced W-byte loads per lane (W = 4, 8, 16:
// the load widths of the DUGKS passes: FP32 g~ / h~ values, (g, h) FP32 face-value pairs,
// 16-byte record loads) and writes one float per work-group; k_probe_write<W> writes a
// buffer once with W-byte stores.  The bytes each kernel must move are known exactly, so
// rocprofv3's FETCH_SIZE / WRITE_SIZE readings of them give the counter scale per load width.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

constexpr int kThreads = 256;

template <int W>
struct Vec;
template <>
struct Vec<4> {
    using T = float;
    __device__ static float sum(T v) { return v; }
    __device__ static T make(float x) { return x; }
};
template <>
struct Vec<8> {
    using T = float2;
    __device__ static float sum(T v) { return v.x + v.y; }
    __device__ static T make(float x) { return float2{x, x}; }
};
template <>
struct Vec<16> {
    using T = float4;
    __device__ static float sum(T v) { return (v.x + v.y) + (v.z + v.w); }
    __device__ static T make(float x) { return float4{x, x, x, x}; }
};

template <int W>
__global__ void __launch_bounds__(kThreads)
    k_probe_read(const typename Vec<W>::T *__restrict__ src, size_t n, float *__restrict__ out)
{
    float acc = 0.0f;
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) acc += Vec<W>::sum(src[i]);
    __shared__ float red[kThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <int W>
__global__ void __launch_bounds__(kThreads) k_probe_write(typename Vec<W>::T *__restrict__ dst, size_t n, float x)
{
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) dst[i] = Vec<W>::make(x);
}

template <int W>
void launch(void *buf, size_t bytes, float *out, int grid, hipStream_t st, bool write)
{
    const size_t n = bytes / W;
    using T = typename Vec<W>::T;
    if (write)
        hipLaunchKernelGGL(k_probe_write<W>, dim3(grid), dim3(kThreads), 0, st, (T *)buf, n, 1.0f);
    else
        hipLaunchKernelGGL(k_probe_read<W>, dim3(grid), dim3(kThreads), 0, st, (const T *)buf, n, out);
}

}  // namespace

extern "C" {

// One streaming read and one streaming write of `bytes` bytes (a multiple of 16) with
// `width`-byte accesses per lane (4, 8 or 16), each timed with HIP events after one untimed
// launch of the same kernel.  read_ms / write_ms receive the timed launch durations.
// Returns 0 on success, 1 for bad arguments, 2 for a HIP error.
int dugks_probe_stream(size_t bytes, int width, double *read_ms, double *write_ms)
{
    if (bytes == 0 || bytes % 16 != 0 || (width != 4 && width != 8 && width != 16)) return 1;
    void *buf = nullptr;
    float *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    int rc = 2;
    const int grid = 8192;  // 32 work-groups per CU, grid-stride loops
    if (hipMalloc(&buf, bytes) != hipSuccess || hipMalloc((void **)&out, grid * sizeof(float)) != hipSuccess) goto done;
    if (hipMemset(buf, 0, bytes) != hipSuccess) goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess)
        goto done;
    for (int rep = 0; rep < 2; ++rep) {
        hipEventRecord(e0, 0);
        if (width == 4) launch<4>(buf, bytes, out, grid, 0, false);
        if (width == 8) launch<8>(buf, bytes, out, grid, 0, false);
        if (width == 16) launch<16>(buf, bytes, out, grid, 0, false);
        hipEventRecord(e1, 0);
        if (width == 4) launch<4>(buf, bytes, out, grid, 0, true);
        if (width == 8) launch<8>(buf, bytes, out, grid, 0, true);
        if (width == 16) launch<16>(buf, bytes, out, grid, 0, true);
        hipEventRecord(e2, 0);
        if (hipEventSynchronize(e2) != hipSuccess || hipGetLastError() != hipSuccess) goto done;
    }
    {
        float a = 0.f, b = 0.f;
        hipEventElapsedTime(&a, e0, e1);
        hipEventElapsedTime(&b, e1, e2);
        if (read_ms) *read_ms = a;
        if (write_ms) *write_ms = b;
    }
    rc = 0;
done:
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e2) hipEventDestroy(e2);
    if (buf) hipFree(buf);
    if (out) hipFree(out);
    return rc;
}

}  // extern "C"
