// Register DFT helpers of the offset-axis FFT kernels (fft2.hip, fft3.hip).
#pragma once
#include "passes.hpp"

namespace adrt {
namespace fftreg {

template <typename T>
__device__ __forceinline__ cplx<T> tw_at(const cplx<T>* tw, int m, int L) {
    // W_L^m for 0 <= m < L from the half table (W^(m + L/2) = -W^m)
    const int h = L >> 1;
    cplx<T> w = tw[m & (h - 1)];
    if (m & h) w = mk(-w.x, -w.y);
    return w;
}

template <typename T, bool INV>
__device__ __forceinline__ cplx<T> rot(cplx<T> w, cplx<T> v) {
    return INV ? cmulc(w, v) : mul(w, v);  // forward W = exp(-i..), inverse conj(W)
}

// In-place radix-2 decimation-in-frequency DFT of M points in registers: result index k sits at position
// bitrev_log2M(k).  w32[m] = W_32^m (m < 16) in the transform's sign convention (conjugated by rot for INV).
template <typename T, int M, bool INV>
__device__ __forceinline__ void dft_dif(cplx<T> (&v)[M], const cplx<T>* w32) {
#pragma unroll
    for (int half = M / 2; half >= 1; half >>= 1) {
#pragma unroll
        for (int a0 = 0; a0 < M; a0 += 2 * half) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const cplx<T> x0 = v[a0 + j], x1 = v[a0 + j + half];
                v[a0 + j] = add(x0, x1);
                const cplx<T> d = sub(x0, x1);
                // W_{2 half}^j = W_32^(j * 16 / half)
                v[a0 + j + half] = (j == 0) ? d : rot<T, INV>(w32[j * (16 / half)], d);
            }
        }
    }
}

template <int M>
__device__ __forceinline__ constexpr int brev_c(int x) {
    int r = 0;
    for (int b = 1; b < M; b <<= 1) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

__device__ __forceinline__ double image_strip2(const double* x, int N, int q, int h, int col) {
    switch (q) {
        case 0: return x[(size_t)col * N + (N - 1 - h)];
        case 1: return x[(size_t)(N - 1 - h) * N + col];
        case 2: return x[(size_t)h * N + col];
        default: return x[(size_t)(N - 1 - col) * N + (N - 1 - h)];
    }
}

}  // namespace fftreg
}  // namespace adrt
