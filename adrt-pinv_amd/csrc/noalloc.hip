// Forward ADRT without a caller workspace: adrt_forward_noalloc, for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// adrt_forward ping-pongs its levels between the output and a workspace of adrt_work_bytes_fwd bytes.  This variant
// runs the same level kernels with that workspace in a file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES.
// Each GPU the library runs on holds its own copy of the array until the process ends; host code never uses the
// array's name as an address (on the host it is a stand-in symbol): the device address comes from
// hipGetSymbolAddress, once per device, with that device current, and is kept per device.  The sinograms run in
// chunks that fit the array, so the size limit is per sinogram: every N up to 2048 in every dtype.
#include <mutex>

#include "../../include/adrt_hip.h"
#include "passes.hpp"

namespace adrt {

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)ADRT_NOALLOC_SCRATCH_MIB << 20;
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on the current device, looked up on the first call for that device
int scratch_on_current_device(void** out) {
    int dev = 0;
    ADRT_HIP(hipGetDevice(&dev));
    ADRT_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "adrt_forward_noalloc: device index beyond the table");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ADRT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ADRT_REQUIRE(p, "adrt_forward_noalloc: null scratch address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OK;
}

size_t elem_bytes(int dtype) { return dtype == 2 ? 8 : 4; }

}  // namespace

// Workspace bytes the forward ADRT of one sinogram writes: levels m < n with n - m odd land in the workspace (the
// others in the output), so it holds the largest of them, level n - 1: 4 (N + N/2 - 1) N entries (none for n = 1).
size_t forward_scratch_per_sinogram(int n, int dtype) {
    const size_t N = (size_t)1 << n;
    size_t most = 0;
    for (int m = 1; m < n; ++m)
        if ((n - m) % 2 == 1) {
            const size_t lv = 4 * (N + ((size_t)2 << (m - 1)) - 1) * N;
            if (lv > most) most = lv;
        }
    return most * elem_bytes(dtype);
}

}  // namespace adrt

using namespace adrt;

extern "C" {

unsigned long long adrt_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int adrt_forward_noalloc(int n, int dtype, const void* d_x, void* d_b, int S, void* stream) {
    ADRT_REQUIRE(n >= 1 && n <= 13, "adrt_forward_noalloc: need 1 <= n <= 13");
    ADRT_REQUIRE(dtype >= 0 && dtype <= 2, "adrt_forward_noalloc: dtype must be 0 (int32), 1 (float) or 2 (double)");
    ADRT_REQUIRE(S >= 0, "adrt_forward_noalloc: S must be >= 0");
    const size_t per = forward_scratch_per_sinogram(n, dtype);
    if (per > NOALLOC_SCRATCH_BYTES) {
        set_error("adrt_forward_noalloc: N = " + std::to_string(1 << n) + " needs " + std::to_string(per) +
                  " bytes of scratch per sinogram, more than the " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                  " bytes of its static scratch; use adrt_forward with a workspace");
        return E_INVALID;
    }
    if (S == 0) return OK;
    ADRT_REQUIRE(d_x && d_b, "adrt_forward_noalloc: null buffer");
    void* work = nullptr;
    ADRT_TRY(scratch_on_current_device(&work));
    const long long N = 1LL << n;
    const size_t el = elem_bytes(dtype);
    const size_t x_bytes = (size_t)(N * N) * el, b_bytes = (size_t)(4 * (2 * N - 1) * N) * el;
    const int chunk = per == 0 ? S : (int)((NOALLOC_SCRATCH_BYTES / per) < (size_t)S ? NOALLOC_SCRATCH_BYTES / per : S);
    // the chunks run in turn on the one stream, so they share the scratch without a race
    for (int s0 = 0; s0 < S; s0 += chunk) {
        const int c = (S - s0) < chunk ? (S - s0) : chunk;
        ADRT_TRY(launch_adrt_forward(n, dtype, (const char*)d_x + (size_t)s0 * x_bytes, (char*)d_b + (size_t)s0 * b_bytes,
                                     work, c, (hipStream_t)stream));
    }
    return OK;
}

}  // extern "C"
