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
