// Runtime services of the C-ABI: memory, copies, fills, streams, events and
// the per-thread error text.  They stand in for the Warp array/runtime calls
// the reference's projection path makes (wp.empty/zeros/copy/from_numpy,
// .numpy(), .zero_(), .fill_(), wp.ScopedTimer; const_param.py:114-150).
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

static thread_local char g_err[1024] = "";

void flub_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

FLUB_API const char *flub_last_error(void) { return g_err; }

FLUB_API int flub_device_count(int *count)
{
    FLUB_REQUIRE(count, FLUB_EINVAL, "flub_device_count: null output");
    FLUB_CHECK_HIP(hipGetDeviceCount(count));
    return FLUB_OK;
}

FLUB_API int flub_malloc(void **ptr, size_t bytes)
{
    FLUB_REQUIRE(ptr, FLUB_EINVAL, "flub_malloc: null output");
    *ptr = nullptr;
    if (bytes == 0) return FLUB_OK;
    FLUB_CHECK_HIP(hipMalloc(ptr, bytes));
    return FLUB_OK;
}

FLUB_API int flub_free(void *ptr)
{
    if (ptr) FLUB_CHECK_HIP(hipFree(ptr));
    return FLUB_OK;
}

FLUB_API int flub_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return FLUB_OK;
    FLUB_REQUIRE(dst && src, FLUB_EINVAL, "flub_memcpy_h2d: null pointer");
    FLUB_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    FLUB_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
    return FLUB_OK;
}

FLUB_API int flub_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return FLUB_OK;
    FLUB_REQUIRE(dst && src, FLUB_EINVAL, "flub_memcpy_d2h: null pointer");
    FLUB_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    FLUB_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
    return FLUB_OK;
}

FLUB_API int flub_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return FLUB_OK;
    FLUB_REQUIRE(dst && src, FLUB_EINVAL, "flub_memcpy_d2d: null pointer");
    FLUB_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return FLUB_OK;
}

FLUB_API int flub_memset(void *ptr, int byte, size_t bytes, void *stream)
{
    if (!bytes) return FLUB_OK;
    FLUB_REQUIRE(ptr, FLUB_EINVAL, "flub_memset: null pointer");
    FLUB_CHECK_HIP(hipMemsetAsync(ptr, byte, bytes, as_stream(stream)));
    return FLUB_OK;
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, T v, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) p[t] = v;
}

template <typename T>
static int fill_impl(T *ptr, T value, size_t n, void *stream)
{
    if (!n) return FLUB_OK;
    FLUB_REQUIRE(ptr, FLUB_EINVAL, "fill: null pointer");
    unsigned nb = blocks_for(n, 256);
    if (nb > 65536) nb = 65536;
    k_fill<T><<<nb, 256, 0, as_stream(stream)>>>(ptr, value, n);
    FLUB_CHECK_LAUNCH();
    return FLUB_OK;
}

FLUB_API int flub_fill_f32(float *ptr, float value, size_t n, void *stream)
{
    return fill_impl<float>(ptr, value, n, stream);
}

FLUB_API int flub_fill_i32(int32_t *ptr, int32_t value, size_t n, void *stream)
{
    return fill_impl<int32_t>(ptr, value, n, stream);
}

FLUB_API int flub_stream_create(void **stream)
{
    FLUB_REQUIRE(stream, FLUB_EINVAL, "flub_stream_create: null output");
    hipStream_t s;
    FLUB_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return FLUB_OK;
}

FLUB_API int flub_stream_destroy(void *stream)
{
    if (stream) FLUB_CHECK_HIP(hipStreamDestroy(as_stream(stream)));
    return FLUB_OK;
}

FLUB_API int flub_stream_synchronize(void *stream)
{
    FLUB_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
    return FLUB_OK;
}

FLUB_API int flub_device_synchronize(void)
{
    FLUB_CHECK_HIP(hipDeviceSynchronize());
    return FLUB_OK;
}

FLUB_API int flub_event_create(void **event)
{
    FLUB_REQUIRE(event, FLUB_EINVAL, "flub_event_create: null output");
    hipEvent_t e;
    FLUB_CHECK_HIP(hipEventCreate(&e));
    *event = e;
    return FLUB_OK;
}

FLUB_API int flub_event_destroy(void *event)
{
    if (event) FLUB_CHECK_HIP(hipEventDestroy((hipEvent_t)event));
    return FLUB_OK;
}

FLUB_API int flub_event_record(void *event, void *stream)
{
    FLUB_REQUIRE(event, FLUB_EINVAL, "flub_event_record: null event");
    FLUB_CHECK_HIP(hipEventRecord((hipEvent_t)event, as_stream(stream)));
    return FLUB_OK;
}

FLUB_API int flub_event_elapsed_ms(void *start, void *stop, float *ms)
{
    FLUB_REQUIRE(start && stop && ms, FLUB_EINVAL, "flub_event_elapsed_ms: null argument");
    FLUB_CHECK_HIP(hipEventSynchronize((hipEvent_t)stop));
    FLUB_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return FLUB_OK;
}
