// phc_common.hip -- error text, device buffers and library-wide entry points.
#include "phc_common.hpp"

namespace phc {

static thread_local std::string g_last_error = "no error";

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

phc_status fail(phc_status code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

phc_status DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) return PHC_OK;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(e == hipErrorOutOfMemory ? PHC_ERR_NOMEM : PHC_ERR_HIP,
                    "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return PHC_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

} // namespace phc

extern "C" {

const char *phc_last_error(void) { return phc::g_last_error.c_str(); }

const char *phc_version(void) { return "phc 0.1 gfx950"; }

int phc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

phc_status phc_device_synchronize(void)
{
    PHC_HIP(hipDeviceSynchronize());
    return PHC_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Bandwidth calibration: a plain streaming copy of known size, used to scale
// the HBM byte counters (FETCH_SIZE / WRITE_SIZE) of the stencil kernels and
// to state the practical copy bandwidth beside the 8 TB/s peak.
__global__ void phc_copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" phc_status phc_copy_bandwidth(size_t bytes, int iters, double *ms_per_copy)
{
    if (!ms_per_copy || iters < 1 || bytes < 16)
        return phc::fail(PHC_ERR_INVALID, "bad copy-bandwidth arguments");
    const size_t n = bytes / 16;
    phc::DevBuf a, b;
    PHC_TRY(a.alloc(n * 16));
    PHC_TRY(b.alloc(n * 16));
    PHC_HIP(hipMemset(a.ptr, 0, n * 16));
    hipEvent_t e0, e1;
    PHC_HIP(hipEventCreate(&e0));
    PHC_HIP(hipEventCreate(&e1));
    const int blocks = 256 * 8;
    hipLaunchKernelGGL(phc_copy_kernel, dim3(blocks), dim3(256), 0, 0, a.as<float4>(), b.as<float4>(), n);
    PHC_HIP(hipEventRecord(e0, 0));
    for (int it = 0; it < iters; ++it)
        hipLaunchKernelGGL(phc_copy_kernel, dim3(blocks), dim3(256), 0, 0, a.as<float4>(),
                           b.as<float4>(), n);
    PHC_HIP(hipEventRecord(e1, 0));
    PHC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PHC_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return PHC_OK;
}
