// C ABI of the device-built environment importance table (declared in include/ptir_amd.h).
#include "../../include/ptir_amd.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "internal.hpp"

namespace {

// Scratch of ptir_env_table_build_noalloc: FP64 cell weights, row sums and two totals for
// tables of up to kNoAllocMaxCells cells and kNoAllocMaxRows rows (1.08 MB per device, held
// until the process ends).  In host code `scratch` only names the symbol: its device
// address, one per GPU, comes from hipGetSymbolAddress.
constexpr int kNoAllocMaxCells = 256 * 512;
constexpr int kNoAllocMaxRows = 4096;
static __device__ __align__(256) double scratch[kNoAllocMaxCells + kNoAllocMaxRows + 2];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, resolved once per device.
hipError_t noalloc_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

ptir_status env_fail(ptir_status st, const std::string& msg) {
    ptir::set_thread_error(msg);
    return st;
}

ptir_status check_env_args(const float* env, int32_t env_h, int32_t env_w, int32_t tab_h, int32_t tab_w,
                           const float* table) {
    if (!env || !table) return env_fail(PTIR_ERR_INVALID_ARGUMENT, "null environment or table pointer");
    if (env_h <= 0 || env_w <= 0) return env_fail(PTIR_ERR_INVALID_ARGUMENT, "environment size must be positive");
    if (env_h == 6 * env_w)
        return env_fail(PTIR_ERR_UNSUPPORTED, "cube-map environments ([6N,N,4]) have no equirect table");
    if (tab_h <= 0 || tab_w <= 0 || tab_h > 4096 || tab_w > 8192 || (int64_t)tab_h * tab_w > (1 << 24))
        return env_fail(PTIR_ERR_INVALID_ARGUMENT, "table size must be positive, at most 4096 x 8192 and 2^24 cells");
    return PTIR_OK;
}

}  // namespace

extern "C" {

uint64_t ptir_env_table_scratch_bytes(int32_t tab_h, int32_t tab_w) {
    if (tab_h <= 0 || tab_w <= 0) return 0;
    return ptir::env_table_scratch_bytes(tab_h, tab_w);
}

ptir_status ptir_env_table_build(const float* environment, int32_t env_h, int32_t env_w, int32_t tab_h, int32_t tab_w,
                                 float* table, void* stream) {
    ptir_status st = check_env_args(environment, env_h, env_w, tab_h, tab_w, table);
    if (st != PTIR_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* scratch = nullptr;
    // stream-ordered scratch: no host synchronisation
    hipError_t e = hipMallocAsync(&scratch, ptir::env_table_scratch_bytes(tab_h, tab_w), s);
    if (e != hipSuccess) return env_fail(PTIR_ERR_OUT_OF_MEMORY, std::string("scratch allocation: ") + hipGetErrorString(e));
    e = ptir::env_table_build(environment, env_h, env_w, tab_h, tab_w, table, scratch, s);
    const hipError_t f = hipFreeAsync(scratch, s);
    if (e == hipSuccess) e = f;
    if (e != hipSuccess) return env_fail(PTIR_ERR_HIP, std::string("environment table build: ") + hipGetErrorString(e));
    return PTIR_OK;
}

ptir_status ptir_env_table_build_noalloc(const float* environment, int32_t env_h, int32_t env_w, int32_t tab_h,
                                         int32_t tab_w, float* table, void* stream) {
    ptir_status st = check_env_args(environment, env_h, env_w, tab_h, tab_w, table);
    if (st != PTIR_OK) return st;
    if ((int64_t)tab_h * tab_w > kNoAllocMaxCells || tab_h > kNoAllocMaxRows)
        return env_fail(PTIR_ERR_INVALID_ARGUMENT,
                        "ptir_env_table_build_noalloc takes tables of at most " + std::to_string(kNoAllocMaxCells) +
                            " cells (" + std::to_string(tab_h) + " x " + std::to_string(tab_w) +
                            " requested); use ptir_env_table_build");
    void* ptr = nullptr;
    hipError_t e = noalloc_scratch(&ptr);
    if (e != hipSuccess) return env_fail(PTIR_ERR_HIP, std::string("resolving the static scratch: ") + hipGetErrorString(e));
    // every scratch word is written by the build before it is read: nothing to clear
    e = ptir::env_table_build(environment, env_h, env_w, tab_h, tab_w, table, ptr, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return env_fail(PTIR_ERR_HIP, std::string("environment table build: ") + hipGetErrorString(e));
    return PTIR_OK;
}

}  // extern "C"
