// nova-mi355x: per-device stream and device-memory helpers.
#include <cstdio>
#include <mutex>

#include <nova/amd/nova_amd.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {
constexpr int kMaxDevices = 64;
hipStream_t g_stream[kMaxDevices];
std::once_flag g_once[kMaxDevices];
} // namespace

hipStream_t stream()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return nullptr;
    std::call_once(g_once[dev], [dev]() {
        hipStream_t s = nullptr;
        if (hipStreamCreate(&s) != hipSuccess)
            s = nullptr;
        g_stream[dev] = s;
    });
    return g_stream[dev];
}

hipStream_t capture_stream()
{
    static hipStream_t cs[kMaxDevices];
    static std::once_flag once[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return nullptr;
    std::call_once(once[dev], [dev]() {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
            s = nullptr;
        cs[dev] = s;
    });
    return cs[dev];
}

std::recursive_mutex &api_mutex()
{
    static std::recursive_mutex mu;
    return mu;
}

int hip_fail(hipError_t e, const char *fn, const char *msg, int code)
{
    if (e == hipSuccess)
        return NOVA_SUCCESS;
    std::fprintf(stderr, "Error[%s]: %s (%s)\n", fn, msg, hipGetErrorString(e));
    return code;
}

int device_alloc(void **p, size_t bytes, const char *fn, const char *what)
{
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 4);
    if (e != hipSuccess) {
        *p = nullptr;
        std::fprintf(stderr, "Error[%s]: Failed to allocate device-side memory for %s (%s)\n",
                     fn, what, hipGetErrorString(e));
        return NOVA_ERROR_DEVICE_MALLOC;
    }
    registry_track(*p);
    return NOVA_SUCCESS;
}

int device_free(void *p, const char *fn, const char *what)
{
    if (p == nullptr)
        return NOVA_SUCCESS;
    registry_untrack(p);
    hipError_t e = hipFree(p);
    if (e != hipSuccess) {
        std::fprintf(stderr, "Error[%s]: Failed to free device-side memory for %s (%s)\n",
                     fn, what, hipGetErrorString(e));
        return NOVA_ERROR_DEVICE_FREE;
    }
    return NOVA_SUCCESS;
}

int device_free_if_owned(void *p, const char *fn, const char *what)
{
    if (p == nullptr || !registry_untrack(p))
        return NOVA_SUCCESS;
    registry_track(p);
    return device_free(p, fn, what);
}

long long first_bad_successor(const int *S, size_t rows, unsigned len, unsigned limit)
{
    for (size_t r = 0; r < rows; r++) {
        const int *row = S + r * len;
        for (unsigned i = 0; i < len; i++) {
            const int sp = row[i];
            if (sp < 0)
                break;
            if ((unsigned)sp >= limit)
                return (long long)(r * len + i);
        }
    }
    return -1;
}

} // namespace nova_amd

namespace nova {

int nova_amd_build_info(char *buf, unsigned int len)
{
    if (buf == nullptr || len == 0)
        return NOVA_ERROR_INVALID_DATA;
    static const char info[] =
        "nova-mi355x; target gfx950 (wave64); kernels: vi_sweep_tiled<E,G>, vi_sweep_plain";
    std::snprintf(buf, len, "%s", info);
    return NOVA_SUCCESS;
}

} // namespace nova
