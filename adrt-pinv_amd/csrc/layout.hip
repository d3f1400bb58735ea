// Sinogram layout permutation between the adrt package's (4, 2N-1, N) array and this library's (the reference's
// pkg_to_core / core_to_pkg, skimage_adrt_pkg.py:38-51): out[s][i] = in[s][src_of[i]], or 0 where src_of[i] < 0 (an
// entry with no counterpart, i.e. a line that misses the image).  The map itself comes from the host (fingerprints of
// both transforms); this is the per-sinogram gather.
//
// HBM-bound: 8 bytes written and 8 read per entry and sinogram, plus the 4-byte map read once per entry for all S
// sinograms (each thread keeps its map entry in a register and walks the batch).  Writes are coalesced; reads follow
// the permutation, which for the package's quadrant flips and transposes stays within few cache lines per wave.
#include "common.hpp"

namespace adrt {
namespace {

constexpr int LAYOUT_NT = 256;

__global__ __launch_bounds__(LAYOUT_NT) void layout_permute_kernel(long long M, const int* __restrict__ src_of,
                                                                   const double* __restrict__ in,
                                                                   double* __restrict__ out, int S) {
    const long long i = (long long)blockIdx.x * LAYOUT_NT + threadIdx.x;
    if (i >= M) return;
    const int j = src_of[i];
    int s = 0;
    for (; s + 4 <= S; s += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = j >= 0 ? in[(long long)(s + u) * M + j] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u) out[(long long)(s + u) * M + i] = v[u];
    }
    for (; s < S; ++s) out[(long long)s * M + i] = j >= 0 ? in[(long long)s * M + j] : 0.0;
}

}  // namespace
}  // namespace adrt

using namespace adrt;

extern "C" {

int adrt_layout_permute(int n, const int* d_src_of, const double* d_in, double* d_out, int S, void* stream) {
    ADRT_REQUIRE(n >= 1 && n <= 13, "layout_permute: need 1 <= n <= 13");
    ADRT_REQUIRE(S >= 0, "layout_permute: S must be >= 0");
    if (S == 0) return OK;
    ADRT_REQUIRE(d_src_of && d_in && d_out, "layout_permute: null buffer");
    ADRT_REQUIRE((const void*)d_in != (const void*)d_out, "layout_permute: d_out must not alias d_in");
    const long long N = 1LL << n, M = 4 * (2 * N - 1) * N;
    const unsigned grid = (unsigned)((M + LAYOUT_NT - 1) / LAYOUT_NT);
    hipLaunchKernelGGL(layout_permute_kernel, dim3(grid), dim3(LAYOUT_NT), 0, (hipStream_t)stream, M, d_src_of, d_in,
                       d_out, S);
    return launch_check("layout_permute_kernel");
}

}  // extern "C"
