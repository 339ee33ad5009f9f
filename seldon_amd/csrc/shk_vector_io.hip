// shk_vector_io.hip -- Seldon's binary vector file for device vectors
// (Vector<VectFull>::Write, vector/Vector.cxx:846-874, and ::Read,
// :1060-1091): the length as an int when with_size, then the elements in
// their binary form.  Without the length, Read takes as many elements as the
// vector already holds.
#include <cstdio>
#include <string>
#include <vector>

#include "shk_internal.hpp"

using namespace shk;

extern "C" {

int shk_vec_write(shk_ctx ctx, shk_dtype dt, int64_t n, const void *dx, const char *path, int with_size)
{
    SHK_CHECK(ctx && path, SHK_ERR_ARG, "vec write: null argument");
    SHK_CHECK(dt == SHK_F64 || dt == SHK_F32, SHK_ERR_ARG, "vec write: bad dtype");
    SHK_CHECK(n >= 0 && n <= 2147483647, SHK_ERR_ARG, "vec write: the length must fit an int");
    const size_t bytes = dsize(dt) * (size_t)n;
    std::vector<char> h(bytes);
    if (n > 0) {
        SHK_TRY(check_device_ptr(dx, "x"));
        SHK_HIP(hipMemcpyAsync(h.data(), dx, bytes, hipMemcpyDeviceToHost, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
    }
    FILE *f = std::fopen(path, "wb");
    SHK_CHECK(f, SHK_ERR_ARG, std::string("vec write: cannot open ") + path);
    const int m = (int)n;
    bool ok = !with_size || std::fwrite(&m, sizeof(int), 1, f) == 1;
    ok = ok && (bytes == 0 || std::fwrite(h.data(), 1, bytes, f) == bytes);
    ok = (std::fclose(f) == 0) && ok;
    SHK_CHECK(ok, SHK_ERR_ARG, std::string("vec write: write failed: ") + path);
    return SHK_OK;
}

int shk_vec_read(shk_ctx ctx, shk_dtype dt, const char *path, int with_size, int64_t n_cap, void *dx, int64_t *n)
{
    SHK_CHECK(ctx && path && n, SHK_ERR_ARG, "vec read: null argument");
    SHK_CHECK(dt == SHK_F64 || dt == SHK_F32, SHK_ERR_ARG, "vec read: bad dtype");
    SHK_CHECK(n_cap >= 0, SHK_ERR_ARG, "vec read: negative capacity");
    FILE *f = std::fopen(path, "rb");
    SHK_CHECK(f, SHK_ERR_ARG, std::string("vec read: cannot open ") + path);
    int64_t len = n_cap;
    if (with_size) {
        int m = -1;
        const bool ok = std::fread(&m, sizeof(int), 1, f) == 1 && m >= 0;
        if (!ok) std::fclose(f);
        SHK_CHECK(ok, SHK_ERR_ARG, std::string("vec read: no length at the start of ") + path);
        len = m;
    }
    *n = len;
    if (!dx) {  // the length only
        std::fclose(f);
        return SHK_OK;
    }
    if (len > n_cap) std::fclose(f);
    SHK_CHECK(len <= n_cap, SHK_ERR_DIM, "vec read: the file's vector is longer than the destination");
    const size_t bytes = dsize(dt) * (size_t)len;
    std::vector<char> h(bytes);
    const bool ok = bytes == 0 || std::fread(h.data(), 1, bytes, f) == bytes;
    std::fclose(f);
    SHK_CHECK(ok, SHK_ERR_ARG, std::string("vec read: fewer elements than expected in ") + path);
    if (len > 0) {
        SHK_TRY(check_device_ptr(dx, "x"));
        SHK_HIP(hipMemcpyAsync(dx, h.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
        SHK_HIP(hipStreamSynchronize(ctx->stream));
    }
    return SHK_OK;
}

}  // extern "C"
