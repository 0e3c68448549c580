// C ABI of the neural radiance cache (declared in include/ptir_amd.h).  No counterpart in
// the reference, which has no radiance cache.
#include <hip/hip_runtime.h>

#include <new>
#include <string>

#include "../../include/ptir_amd.h"
#include "internal.hpp"

struct ptir_cache {
    ptir_cache_options opts;
    ptir::NrcBuffers buf;
    std::string err;
};

namespace {

thread_local std::string g_cache_create_error;

ptir_status cfail(ptir_cache* c, ptir_status st, const std::string& msg) {
    if (c) c->err = msg;
    else g_cache_create_error = msg;
    return st;
}

ptir_status chip(ptir_cache* c, hipError_t e, const char* where) {
    return cfail(c, PTIR_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

ptir::nrc::AdamParams adam(const ptir_cache_options& o) {
    return ptir::nrc::AdamParams{o.learning_rate, o.beta1, o.beta2, o.epsilon};
}

}  // namespace

extern "C" {

ptir_status ptir_cache_create(const ptir_cache_options* o, ptir_cache** out, void* stream) {
    if (!o || !out) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null options or output handle");
    *out = nullptr;
    // scratch grows with max_batch: 68 KB of per-tile weight gradients per 64 records, so cap
    // it at 2^20 records (1.1 GB)
    if (o->max_batch == 0 || o->max_batch > (1ull << 20))
        return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "max_batch must be in [1, 2^20]");
    if (!(o->learning_rate >= 0.f) || !(o->beta1 >= 0.f && o->beta1 < 1.f) || !(o->beta2 >= 0.f && o->beta2 < 1.f) ||
        !(o->epsilon > 0.f))
        return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "Adam needs lr >= 0, beta1 and beta2 in [0,1), epsilon > 0");
    ptir_cache* c = new (std::nothrow) ptir_cache();
    if (!c) return cfail(nullptr, PTIR_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->opts = *o;
    hipError_t e = ptir::nrc_alloc(c->buf, o->max_batch, o->seed, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        const std::string msg = std::string("cache allocation: ") + hipGetErrorString(e);
        ptir::nrc_free(c->buf);
        delete c;
        return cfail(nullptr, e == hipErrorOutOfMemory ? PTIR_ERR_OUT_OF_MEMORY : PTIR_ERR_HIP, msg);
    }
    *out = c;
    return PTIR_OK;
}

ptir_status ptir_cache_destroy(ptir_cache* c) {
    if (!c) return PTIR_OK;
    ptir::nrc_free(c->buf);
    delete c;
    return PTIR_OK;
}

uint64_t ptir_cache_num_weights(void) { return ptir::nrc::kParams; }

uint64_t ptir_cache_table_floats(void) { return ptir::nrc_table_floats(); }

ptir_status ptir_cache_read_params(ptir_cache* c, float* weights, float* table, void* stream) {
    if (!c) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null cache");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (weights) e = hipMemcpyAsync(weights, c->buf.w, ptir::nrc::kParams * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && table)
        e = hipMemcpyAsync(table, c->buf.table, ptir::nrc_table_floats() * sizeof(float), hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? PTIR_OK : chip(c, e, "reading cache parameters");
}

ptir_status ptir_cache_write_params(ptir_cache* c, const float* weights, const float* table, void* stream) {
    if (!c) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null cache");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (weights) e = hipMemcpyAsync(c->buf.w, weights, ptir::nrc::kParams * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && table)
        e = hipMemcpyAsync(c->buf.table, table, ptir::nrc_table_floats() * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = ptir::nrc_pack(c->buf, s);
    return e == hipSuccess ? PTIR_OK : chip(c, e, "writing cache parameters");
}

ptir_status ptir_cache_read_grads(ptir_cache* c, float* weight_grad, float* table_grad, float* loss, void* stream) {
    if (!c) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null cache");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (weight_grad)
        e = hipMemcpyAsync(weight_grad, c->buf.dw, ptir::nrc::kParams * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess && table_grad)
        e = hipMemcpyAsync(table_grad, c->buf.dtable, ptir::nrc_table_floats() * sizeof(float), hipMemcpyDeviceToDevice,
                           s);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, c->buf.loss, sizeof(float), hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? PTIR_OK : chip(c, e, "reading cache gradients");
}

ptir_status ptir_cache_query(ptir_cache* c, const float* records, uint64_t n, const float* box, float* pbr,
                             float* comp, float* pred, void* stream) {
    if (!c) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null cache");
    if (n && (!records || !box)) return cfail(c, PTIR_ERR_INVALID_ARGUMENT, "null records or scene box");
    if (n > (1ull << 34)) return cfail(c, PTIR_ERR_INVALID_ARGUMENT, "too many records");
    hipError_t e = ptir::nrc_query(c->buf, records, n, box, pbr, comp, pred, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? PTIR_OK : chip(c, e, "cache query");
}

ptir_status ptir_cache_backprop(ptir_cache* c, const float* records, uint64_t n, const float* box, void* stream) {
    if (!c) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null cache");
    if (n > c->buf.max_batch)
        return cfail(c, PTIR_ERR_INVALID_ARGUMENT, "training batch of " + std::to_string(n) + " records exceeds max_batch " +
                                                       std::to_string(c->buf.max_batch));
    if (n && (!records || !box)) return cfail(c, PTIR_ERR_INVALID_ARGUMENT, "null records or scene box");
    hipError_t e = ptir::nrc_backprop(c->buf, records, n, box, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? PTIR_OK : chip(c, e, "cache backprop");
}

ptir_status ptir_cache_step(ptir_cache* c, void* stream) {
    if (!c) return cfail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null cache");
    hipError_t e = ptir::nrc_step(c->buf, adam(c->opts), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? PTIR_OK : chip(c, e, "cache step");
}

ptir_status ptir_cache_train(ptir_cache* c, const float* records, uint64_t n, const float* box, void* stream) {
    ptir_status st = ptir_cache_backprop(c, records, n, box, stream);
    return st == PTIR_OK ? ptir_cache_step(c, stream) : st;
}

const char* ptir_cache_last_error(const ptir_cache* c) { return c ? c->err.c_str() : g_cache_create_error.c_str(); }

}  // extern "C"
