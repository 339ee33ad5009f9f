// shk_context.hip -- context, errors and device memory of the shk library.
#include <cstring>

#include "shk_internal.hpp"

namespace shk {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int ensure_red(shk_ctx ctx, size_t bytes)
{
    if (ctx->red.cap_bytes >= bytes) return SHK_OK;
    if (ctx->red.partials) SHK_HIP(hipFree(ctx->red.partials));
    size_t cap = bytes < (1u << 20) ? (1u << 20) : bytes;
    SHK_HIP(hipMalloc(&ctx->red.partials, cap));
    ctx->red.cap_bytes = cap;
    return SHK_OK;
}

int ensure_ws(shk_ctx ctx, size_t bytes)
{
    if (ctx->ws_cap >= bytes) return SHK_OK;
    if (ctx->ws) {
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        SHK_HIP(hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_cap = 0;
    }
    hipError_t e = hipMalloc(&ctx->ws, bytes);
    if (e != hipSuccess)
        return fail(SHK_ERR_NOMEM, std::string("solver workspace: ") + hipGetErrorString(e));
    ctx->ws_cap = bytes;
    return SHK_OK;
}

int ensure_ws2(shk_ctx ctx, size_t bytes)
{
    if (ctx->ws2_cap >= bytes) return SHK_OK;
    if (ctx->ws2) {
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        SHK_HIP(hipFree(ctx->ws2));
        ctx->ws2 = nullptr;
        ctx->ws2_cap = 0;
    }
    hipError_t e = hipMalloc(&ctx->ws2, bytes);
    if (e != hipSuccess)
        return fail(SHK_ERR_NOMEM, std::string("refinement workspace: ") + hipGetErrorString(e));
    ctx->ws2_cap = bytes;
    return SHK_OK;
}

int check_device_ptr(const void *p, const char *what)
{
    if (!p) return fail(SHK_ERR_ARG, std::string("null device pointer: ") + what);
    return SHK_OK;
}

}  // namespace shk

using namespace shk;

extern "C" {

const char *shk_last_error(void) { return g_last_error.c_str(); }

const char *shk_version(void) { return "shk 0.1 (gfx950)"; }

int shk_ctx_create(int device, shk_ctx *out)
{
    SHK_CHECK(out, SHK_ERR_ARG, "shk_ctx_create: null out");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(SHK_ERR_NODEV, "shk_ctx_create: no HIP device visible");
    SHK_CHECK(device >= 0 && device < count, SHK_ERR_ARG, "shk_ctx_create: bad device index");
    SHK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SHK_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SHK_ERR_NODEV, std::string("shk_ctx_create: built for gfx950, device is ") +
                                       prop.gcnArchName);
    shk_ctx c = new shk_ctx_s();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    std::strncpy(c->arch, prop.gcnArchName, sizeof(c->arch) - 1);
    // default priority, as every stream of this library
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SHK_ERR_HIP, "shk_ctx_create: hipStreamCreate failed");
    }
    c->stream = c->own;
    int st = ensure_red(c, 1u << 20);
    if (st != SHK_OK) { delete c; return st; }
    st = noalloc_scratch_init(device);
    if (st != SHK_OK) { delete c; return st; }
    SHK_HIP(hipMalloc(&c->scalars, 4096));
    SHK_HIP(hipMemset(c->scalars, 0, 4096));
    SHK_HIP(hipMalloc(&c->red.ticket, 64 * sizeof(unsigned)));
    SHK_HIP(hipMemset(c->red.ticket, 0, 64 * sizeof(unsigned)));
    c->state_bytes = 1 << 16;
    SHK_HIP(hipMalloc(&c->state, c->state_bytes));
    SHK_HIP(hipMemset(c->state, 0, c->state_bytes));
    SHK_HIP(hipHostMalloc(&c->state_host, c->state_bytes, hipHostMallocDefault));
    SHK_HIP(hipEventCreate(&c->ev0));
    SHK_HIP(hipEventCreate(&c->ev1));
    *out = c;
    return SHK_OK;
}

int shk_ctx_destroy(shk_ctx c)
{
    if (!c) return SHK_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (c->red.partials) hipFree(c->red.partials);
    if (c->red.ticket) hipFree(c->red.ticket);
    if (c->scalars) hipFree(c->scalars);
    if (c->ws) hipFree(c->ws);
    if (c->ws2) hipFree(c->ws2);
    if (c->state) hipFree(c->state);
    if (c->state_host) hipHostFree(c->state_host);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->own) hipStreamDestroy(c->own);
    delete c;
    return SHK_OK;
}

int shk_ctx_set_stream(shk_ctx c, void *s)
{
    SHK_CHECK(c, SHK_ERR_ARG, "null ctx");
    SHK_HIP(hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)s;  // verbatim: NULL is the default (null) stream
    return SHK_OK;
}

int shk_ctx_synchronize(shk_ctx c)
{
    SHK_CHECK(c, SHK_ERR_ARG, "null ctx");
    SHK_HIP(hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_ctx_device_info(shk_ctx c, int *num_cus, int *num_xcds, char *arch, int arch_len)
{
    SHK_CHECK(c, SHK_ERR_ARG, "null ctx");
    if (num_cus) *num_cus = c->num_cus;
    if (num_xcds) *num_xcds = kNumXcd;
    if (arch && arch_len > 0) {
        std::strncpy(arch, c->arch, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return SHK_OK;
}

int shk_malloc(shk_ctx c, size_t bytes, void **p)
{
    SHK_CHECK(c && p, SHK_ERR_ARG, "shk_malloc: null argument");
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(SHK_ERR_NOMEM, std::string("shk_malloc: ") + hipGetErrorString(e));
    return SHK_OK;
}

int shk_free(shk_ctx c, void *p)
{
    SHK_CHECK(c, SHK_ERR_ARG, "shk_free: null ctx");
    if (p) {
        SHK_HIP(hipStreamSynchronize(c->stream));
        SHK_HIP(hipFree(p));
    }
    return SHK_OK;
}

int shk_memcpy_h2d(shk_ctx c, void *dst, const void *src, size_t bytes)
{
    SHK_CHECK(c, SHK_ERR_ARG, "null ctx");
    if (!bytes) return SHK_OK;
    SHK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SHK_HIP(hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_memcpy_d2h(shk_ctx c, void *dst, const void *src, size_t bytes)
{
    SHK_CHECK(c, SHK_ERR_ARG, "null ctx");
    if (!bytes) return SHK_OK;
    SHK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SHK_HIP(hipStreamSynchronize(c->stream));
    return SHK_OK;
}

int shk_memcpy_d2d(shk_ctx c, void *dst, const void *src, size_t bytes)
{
    SHK_CHECK(c, SHK_ERR_ARG, "null ctx");
    if (!bytes) return SHK_OK;
    SHK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return SHK_OK;
}

}  // extern "C"
