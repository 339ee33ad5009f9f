// shk_blas1.hip -- BLAS-1 entry points (Seldon's Functions_Vector.cxx set).
#include "shk_kernels.hpp"

using namespace shk;

namespace {

template <typename T>
int dot_impl(shk_ctx ctx, int64_t n, const T *x, const T *y, const T *u1, T alpha,
             const T *ax, T *ay, T *dres)
{
    // result[0] = sum u0*v, result[1] = sum u1*v with v = y (+ alpha ax)
    const int nb = red_blocks(n);
    SHK_TRY(ensure_red(ctx, sizeof(T) * 2 * nb));
    hipLaunchKernelGGL(k_axpy_dot<T>, dim3(nb), dim3(kBlock), 0, ctx->stream, n, alpha, ax,
                       ay ? ay : (T *)y, x, u1, (T *)ctx->red.partials, (const int *)nullptr);
    grid_finish<T>(ctx->stream, (const T *)ctx->red.partials, nb, nullptr, FinStore<T>{dres});
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

template <typename T>
int fetch_scalar(shk_ctx ctx, const T *d, double *out)
{
    T h = 0;
    SHK_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    SHK_HIP(hipStreamSynchronize(ctx->stream));
    *out = (double)h;
    return SHK_OK;
}

int check_n(shk_ctx ctx, int64_t n, shk_dtype dt)
{
    SHK_CHECK(ctx, SHK_ERR_ARG, "null ctx");
    SHK_CHECK(n >= 0, SHK_ERR_ARG, "negative length");
    SHK_CHECK(dt == SHK_F64 || dt == SHK_F32, SHK_ERR_ARG, "bad dtype");
    return SHK_OK;
}

}  // namespace

extern "C" {

int shk_dot_device(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, const void *dy, void *dres)
{
    SHK_TRY(check_n(ctx, n, dt));
    SHK_TRY(check_device_ptr(dres, "result"));
    if (n > 0) {
        SHK_TRY(check_device_ptr(dx, "x"));
        SHK_TRY(check_device_ptr(dy, "y"));
    }
    // DotProd sums X(i) * Y(i): u0 = x, v = y
    if (dt == SHK_F64)
        return dot_impl<double>(ctx, n, (const double *)dx, (const double *)dy, nullptr, 0.0, nullptr,
                                nullptr, (double *)dres);
    return dot_impl<float>(ctx, n, (const float *)dx, (const float *)dy, nullptr, 0.0f, nullptr, nullptr,
                           (float *)dres);
}

int shk_dot(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, const void *dy, double *result)
{
    SHK_CHECK(result, SHK_ERR_ARG, "null result");
    SHK_TRY(shk_dot_device(ctx, dt, n, dx, dy, ctx ? ctx->scalars : nullptr));
    if (dt == SHK_F64) return fetch_scalar<double>(ctx, (const double *)ctx->scalars, result);
    return fetch_scalar<float>(ctx, (const float *)ctx->scalars, result);
}

int shk_nrm2(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, double *result)
{
    SHK_CHECK(result, SHK_ERR_ARG, "null result");
    SHK_TRY(shk_dot_device(ctx, dt, n, dx, dx, ctx ? ctx->scalars : nullptr));
    if (dt == SHK_F64) {
        double h = 0;
        SHK_HIP(hipMemcpyAsync(&h, ctx->scalars, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        *result = std::sqrt(h);
    } else {
        float h = 0;
        SHK_HIP(hipMemcpyAsync(&h, ctx->scalars, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
        *result = (double)std::sqrt(h);
    }
    return SHK_OK;
}

int shk_axpy(shk_ctx ctx, shk_dtype dt, int64_t n, double alpha, const void *dx, void *dy)
{
    SHK_TRY(check_n(ctx, n, dt));
    if (n == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy, "y"));
    if (dt == SHK_F64) {
        if (alpha == 0.0) return SHK_OK;  // Add skips alpha == 0
        hipLaunchKernelGGL(k_axpy<double>, dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n, alpha,
                           (const double *)dx, (double *)dy);
    } else {
        float a = (float)alpha;
        if (a == 0.0f) return SHK_OK;
        hipLaunchKernelGGL(k_axpy<float>, dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n, a,
                           (const float *)dx, (float *)dy);
    }
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

int shk_scal(shk_ctx ctx, shk_dtype dt, int64_t n, double alpha, void *dx)
{
    SHK_TRY(check_n(ctx, n, dt));
    if (n == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    if (dt == SHK_F64)
        hipLaunchKernelGGL(k_scal<double>, dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n, alpha,
                           (double *)dx);
    else
        hipLaunchKernelGGL(k_scal<float>, dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n,
                           (float)alpha, (float *)dx);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

int shk_copy(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, void *dy)
{
    SHK_TRY(check_n(ctx, n, dt));
    if (n == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    SHK_TRY(check_device_ptr(dy, "y"));
    if (dx == dy) return SHK_OK;
    SHK_HIP(hipMemcpyAsync(dy, dx, n * dsize(dt), hipMemcpyDeviceToDevice, ctx->stream));
    return SHK_OK;
}

int shk_fill(shk_ctx ctx, shk_dtype dt, int64_t n, double value, void *dx)
{
    SHK_TRY(check_n(ctx, n, dt));
    if (n == 0) return SHK_OK;
    SHK_TRY(check_device_ptr(dx, "x"));
    if (dt == SHK_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n, value,
                           (double *)dx, (const int *)nullptr);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n,
                           (float)value, (float *)dx, (const int *)nullptr);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

int shk_axpy_dot(shk_ctx ctx, shk_dtype dt, int64_t n, double alpha, const void *dx, void *dy,
                 const void *dz, double *result)
{
    SHK_TRY(check_n(ctx, n, dt));
    SHK_CHECK(result, SHK_ERR_ARG, "null result");
    if (n > 0) {
        SHK_TRY(check_device_ptr(dx, "x"));
        SHK_TRY(check_device_ptr(dy, "y"));
        SHK_TRY(check_device_ptr(dz, "z"));
    }
    // result = sum z * y after y += alpha x
    if (dt == SHK_F64) {
        SHK_TRY(dot_impl<double>(ctx, n, (const double *)dz, (const double *)dy, nullptr, alpha,
                                 (const double *)dx, (double *)dy, (double *)ctx->scalars));
        return fetch_scalar<double>(ctx, (const double *)ctx->scalars, result);
    }
    SHK_TRY(dot_impl<float>(ctx, n, (const float *)dz, (const float *)dy, nullptr, (float)alpha,
                            (const float *)dx, (float *)dy, (float *)ctx->scalars));
    return fetch_scalar<float>(ctx, (const float *)ctx->scalars, result);
}

int shk_convert(shk_ctx ctx, int64_t n, shk_dtype sdt, const void *src, shk_dtype ddt, void *dst)
{
    SHK_TRY(check_n(ctx, n, sdt));
    if (n == 0) return SHK_OK;
    if (sdt == ddt) return shk_copy(ctx, sdt, n, src, dst);
    if (sdt == SHK_F64)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n,
                           (const double *)src, (float *)dst);
    else
        hipLaunchKernelGGL((k_convert<float, double>), dim3(red_blocks(n)), dim3(kBlock), 0, ctx->stream, n,
                           (const float *)src, (double *)dst);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

}  // extern "C"
