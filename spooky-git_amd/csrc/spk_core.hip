// spk_core.hip -- error text, ABI version, contexts and memory helpers.
#include "spk_internal.hpp"

#include <cstring>

namespace spk {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

spk_ctx ctx_retain(spk_ctx ctx) {
    ctx->refs.fetch_add(1, std::memory_order_relaxed);
    return ctx;
}

void ctx_release(spk_ctx ctx) {
    if (ctx && ctx->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete ctx;
}

int make_batch(spk_ctx ctx, std::initializer_list<const void *> ptrs, Batch *out) {
    memset(out, 0, sizeof(*out));
    for (int i = 0; i < BATCH_ARGS; i++) out->slot[i] = -1;
    if (!ctx->rows) return SPK_OK;
    SPK_REQUIRE((int)ptrs.size() <= BATCH_ARGS, SPK_ERR_STATE, "batched launch: %zu pointer arguments", ptrs.size());
    out->rows = ctx->rows;
    int i = 0;
    for (const void *p : ptrs) {
        if (p != nullptr) {
            const char *c = static_cast<const char *>(p);
            int slot = -1;
            for (int k = 0; k < BOX_SLOTS && slot < 0; k++)
                if (ctx->base[k] && c >= ctx->base[k] && c < ctx->base[k] + ctx->size[k]) slot = k;
            SPK_REQUIRE(slot >= 0, SPK_ERR_STATE, "batched launch: argument %d (%p) is in no array of the batch", i, p);
            out->slot[i] = (signed char)slot;
            out->off[i] = (long long)(c - ctx->base[slot]);
        }
        i++;
    }
    return SPK_OK;
}

}  // namespace spk

extern "C" int spk_abi_version(void) { return SPK_ABI_VERSION; }

extern "C" const char *spk_last_error(void) { return spk::g_err; }

extern "C" size_t spk_params_size(void) { return sizeof(spk_params); }

extern "C" int spk_ctx_create(int device, void *stream, spk_ctx *out) {
    SPK_REQUIRE(out, SPK_ERR_ARG, "spk_ctx_create: null out");
    int n = 0;
    SPK_HIP(hipGetDeviceCount(&n));
    SPK_REQUIRE(device >= 0 && device < n, SPK_ERR_ARG, "spk_ctx_create: device %d of %d", device, n);
    SPK_HIP(hipSetDevice(device));
    SPK_TRY(spk::noalloc_setup(device));
    spk_ctx c = new spk_ctx_s();
    c->device = device;
    c->stream = (hipStream_t)stream;
    *out = c;
    return SPK_OK;
}

extern "C" int spk_ctx_destroy(spk_ctx ctx) {
    spk::ctx_release(ctx);
    return SPK_OK;
}

extern "C" int spk_ctx_sync(spk_ctx ctx) {
    SPK_REQUIRE(ctx, SPK_ERR_ARG, "spk_ctx_sync: null ctx");
    SPK_HIP(hipStreamSynchronize(ctx->stream));
    return SPK_OK;
}

extern "C" int spk_malloc(spk_ctx ctx, size_t bytes, void **out) {
    SPK_REQUIRE(ctx && out, SPK_ERR_ARG, "spk_malloc: null argument");
    SPK_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        spk::set_error("spk_malloc(%zu): %s", bytes, hipGetErrorString(e));
        return SPK_ERR_ALLOC;
    }
    return SPK_OK;
}

extern "C" int spk_free(spk_ctx ctx, void *ptr) {
    (void)ctx;
    if (ptr) SPK_HIP(hipFree(ptr));
    return SPK_OK;
}

extern "C" int spk_memcpy_h2d(spk_ctx ctx, void *dst, const void *src, size_t bytes) {
    SPK_REQUIRE(ctx, SPK_ERR_ARG, "spk_memcpy_h2d: null ctx");
    SPK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SPK_HIP(hipStreamSynchronize(ctx->stream));
    return SPK_OK;
}

extern "C" int spk_memcpy_d2h(spk_ctx ctx, void *dst, const void *src, size_t bytes) {
    SPK_REQUIRE(ctx, SPK_ERR_ARG, "spk_memcpy_d2h: null ctx");
    SPK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SPK_HIP(hipStreamSynchronize(ctx->stream));
    return SPK_OK;
}
