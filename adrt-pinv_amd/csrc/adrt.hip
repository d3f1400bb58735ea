// Forward ADRT and its transpose (adrtcore.py:63-137), one launch per level over all (sinogram, quadrant, row, col).
// Every output entry is (0 + half-0 term) + half-1 term exactly as the reference accumulates, so integer inputs are
// exact and float results are bit-identical to the NumPy levels.  Memory-bound integer/float adds: one thread per
// output entry, coalesced along the slope axis.
#include "passes.hpp"

namespace adrt {

namespace {

constexpr int ADRT_THREADS = 256;

inline unsigned grid_for(long long count) {
    long long g = (count + ADRT_THREADS - 1) / ADRT_THREADS;
    if (g > (1LL << 20)) g = 1LL << 20;
    return (unsigned)(g < 1 ? 1 : g);
}

// value of quadrant q of image x at (i, j): T_q x (adrtcore.py:29-39)
template <typename Ti>
__device__ __forceinline__ Ti quad_image(const Ti* x, int N, int q, int i, int j) {
    switch (q) {
        case 0: return x[(size_t)j * N + (N - 1 - i)];
        case 1: return x[(size_t)(N - 1 - i) * N + j];
        case 2: return x[(size_t)i * N + j];
        default: return x[(size_t)(N - 1 - j) * N + (N - 1 - i)];
    }
}

// Level m: in (S*4, N + w - 1, N) -> out (S*4, N + 2w - 1, N).  For m == 1 the input is the image (S, N, N).
template <typename Tin, typename Tout>
__global__ void adrt_level_kernel(int n, int m, const Tin* in, Tout* out, long long count) {
    const int N = 1 << n, w = 1 << (m - 1);
    const int Hin = N + w - 1, Hout = N + 2 * w - 1;
    for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < count; o += (long long)gridDim.x * blockDim.x) {
        const int col = (int)(o % N);
        const int row = (int)((o / N) % Hout);
        const long long sq = o / ((long long)N * Hout);
        const int h = row - (2 * w - 1);
        const int l = col / (2 * w), j = col % (2 * w);
        const int t = j >> 1, b = j & 1;
        const int c0 = l * 2 * w + t, c1 = c0 + w;
        Tout acc = 0;
        if (m == 1) {
            const int s = (int)(sq >> 2), q = (int)(sq & 3);
            const Tin* x = in + (size_t)s * N * N;
            if (h >= -t) acc = acc + (Tout)quad_image(x, N, q, h, c0);
            const int hs = h + t + b;
            if (hs >= -t && hs < N) acc = acc + (Tout)quad_image(x, N, q, hs, c1);
        } else {
            const Tin* f = in + sq * (long long)Hin * N;
            if (h >= -t) acc = acc + (Tout)f[(size_t)(h + w - 1) * N + c0];
            const int hs = h + t + b;
            if (hs >= -t && hs < N) acc = acc + (Tout)f[(size_t)(hs + w - 1) * N + c1];
        }
        out[o] = acc;
    }
}

// Transpose of level m: g (S*4, N + 2w - 1, N) -> f (S*4, N + w - 1, N) (adrtcore.py:86-101).
template <typename Tin, typename Tout>
__global__ void adrt_level_T_kernel(int n, int m, const Tin* g, Tout* f, long long count) {
    const int N = 1 << n, w = 1 << (m - 1);
    const int Hin = N + w - 1, Hout = N + 2 * w - 1;
    for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < count; o += (long long)gridDim.x * blockDim.x) {
        const int col = (int)(o % N);
        const int row = (int)((o / N) % Hin);
        const long long sq = o / ((long long)N * Hin);
        const int h = row - (w - 1);
        const int l = col / (2 * w), jj = col % (2 * w);
        const int half = jj / w, t = jj % w;
        const int e0 = l * 2 * w + 2 * t, e1 = e0 + 1;
        const Tin* gs = g + sq * (long long)Hout * N;
        Tout acc = 0;
        if (h >= -t) {
            const int ra = h - (half ? t : 0);
            const int rb = ra - half;
            acc = acc + (Tout)gs[(size_t)(ra + 2 * w - 1) * N + e0];
            acc = acc + (Tout)gs[(size_t)(rb + 2 * w - 1) * N + e1];
        }
        f[o] = acc;
    }
}

// x[s][a][b] = (((0 + f0[N-1-b][a]) + f1[N-1-a][b]) + f2[a][b]) + f3[N-1-b][N-1-a]  (sum of T_q^T, adrtcore.py:131-137)
template <typename T>
__global__ void quad_sum_kernel(int n, const T* f, T* x, long long count) {
    const int N = 1 << n;
    const long long NN = (long long)N * N;
    for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < count; o += (long long)gridDim.x * blockDim.x) {
        const long long s = o / NN;
        const int a = (int)((o % NN) / N), b = (int)(o % N);
        const T* f0 = f + s * 4 * NN;
        const T* f1 = f0 + NN;
        const T* f2 = f1 + NN;
        const T* f3 = f2 + NN;
        T acc = 0;
        acc = acc + f0[(size_t)(N - 1 - b) * N + a];
        acc = acc + f1[(size_t)(N - 1 - a) * N + b];
        acc = acc + f2[(size_t)a * N + b];
        acc = acc + f3[(size_t)(N - 1 - b) * N + (N - 1 - a)];
        x[o] = acc;
    }
}

template <typename T>
int forward_impl(int n, const T* x, T* b, T* work, int S, hipStream_t st) {
    const int N = 1 << n;
    const T* in = x;
    for (int m = 1; m <= n; ++m) {
        T* out = ((n - m) % 2 == 0) ? b : work;
        const long long count = (long long)S * 4 * (N + (2LL << (m - 1)) - 1) * N;
        hipLaunchKernelGGL((adrt_level_kernel<T, T>), dim3(grid_for(count)), dim3(ADRT_THREADS), 0, st, n, m, in, out,
                           count);
        ADRT_TRY(launch_check("adrt_level_kernel"));
        in = out;
    }
    return OK;
}

template <typename Tin, typename Ti>
int transpose_impl(int n, const Tin* b, Ti* x, Ti* work, int S, hipStream_t st) {
    const int N = 1 << n;
    const size_t full = (size_t)S * 4 * (2 * N - 1) * N;
    Ti* w0 = work;
    Ti* w1 = work + full;
    // level n reads the (typed) input, later levels ping-pong in the intermediate type
    {
        const long long count = (long long)S * 4 * (N + (1LL << (n - 1)) - 1) * N;
        hipLaunchKernelGGL((adrt_level_T_kernel<Tin, Ti>), dim3(grid_for(count)), dim3(ADRT_THREADS), 0, st, n, n, b,
                           w0, count);
        ADRT_TRY(launch_check("adrt_level_T_kernel"));
    }
    Ti* cur = w0;
    Ti* nxt = w1;
    for (int m = n - 1; m >= 1; --m) {
        const long long count = (long long)S * 4 * (N + (1LL << (m - 1)) - 1) * N;
        hipLaunchKernelGGL((adrt_level_T_kernel<Ti, Ti>), dim3(grid_for(count)), dim3(ADRT_THREADS), 0, st, n, m, cur,
                           nxt, count);
        ADRT_TRY(launch_check("adrt_level_T_kernel"));
        Ti* t = cur;
        cur = nxt;
        nxt = t;
    }
    const long long count = (long long)S * N * N;
    hipLaunchKernelGGL((quad_sum_kernel<Ti>), dim3(grid_for(count)), dim3(ADRT_THREADS), 0, st, n, cur, x, count);
    return launch_check("quad_sum_kernel");
}

}  // namespace

size_t adrt_work_bytes(int n, int dtype, int S, int transpose) {
    const size_t N = (size_t)1 << n;
    const size_t full = (size_t)S * 4 * (2 * N - 1) * N;
    size_t elem = dtype == 0 ? (transpose ? 8 : 4) : (dtype == 1 ? 4 : 8);
    return full * elem * (transpose ? 2 : 1);
}

int launch_adrt_forward(int n, int dtype, const void* x, void* b, void* work, int S, hipStream_t st) {
    ADRT_REQUIRE(n >= 1 && n <= 13, "adrt_forward: need 1 <= n <= 13");
    ADRT_REQUIRE(dtype >= 0 && dtype <= 2, "adrt_forward: dtype must be 0 (int32), 1 (float) or 2 (double)");
    if (S <= 0) return OK;
    ADRT_REQUIRE(x && b && (work || n == 1), "adrt_forward: null buffer");
    if (dtype == 0) return forward_impl<int32_t>(n, (const int32_t*)x, (int32_t*)b, (int32_t*)work, S, st);
    if (dtype == 1) return forward_impl<float>(n, (const float*)x, (float*)b, (float*)work, S, st);
    return forward_impl<double>(n, (const double*)x, (double*)b, (double*)work, S, st);
}

int launch_adrt_transpose(int n, int dtype, const void* b, void* x, void* work, int S, hipStream_t st) {
    ADRT_REQUIRE(n >= 1 && n <= 13, "adrt_transpose: need 1 <= n <= 13");
    ADRT_REQUIRE(dtype >= 0 && dtype <= 2, "adrt_transpose: dtype must be 0 (int32 -> int64), 1 (float) or 2 (double)");
    if (S <= 0) return OK;
    ADRT_REQUIRE(x && b && work, "adrt_transpose: null buffer");
    if (dtype == 0) return transpose_impl<int32_t, int64_t>(n, (const int32_t*)b, (int64_t*)x, (int64_t*)work, S, st);
    if (dtype == 1) return transpose_impl<float, float>(n, (const float*)b, (float*)x, (float*)work, S, st);
    return transpose_impl<double, double>(n, (const double*)b, (double*)x, (double*)work, S, st);
}

}  // namespace adrt
