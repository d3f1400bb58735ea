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

// Mode-0 packing of one pair's two strips for a block's C columns c0 .. c0 + C - 1.  The pair's transposed strip
// (q = 0 for pair 0, q = 3 for pair 1) is C whole image rows: stage_rows copies them into LDS with coalesced loads
// (C rows of N + 1 values, rounded to T as the strip read would), and strip_pair takes that strip's value from there
// and the other strip's x[.][c] from memory, which is coalesced across the block's columns.
template <typename T, int N, int C, int NT>
__device__ __forceinline__ void stage_rows(const double* x, int pair, int c0, T* stg, int t) {
    static_assert((C * N) % NT == 0, "stage_rows: the rows must split evenly over the threads");
    constexpr int PER = C * N / NT;
    const int row0 = pair == 0 ? c0 : N - c0 - C;
    const double* src = x + (size_t)row0 * N;  // C consecutive rows: one contiguous span
    double val[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) val[e] = src[t + NT * e];  // every load in flight before the first store
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int i = t + NT * e, r = i / N, j = i % N;
        stg[r * (N + 1) + j] = (T)val[e];
    }
}
template <typename T>
__device__ __forceinline__ cplx<T> strip_pair(const double* x, int N, int pair, int h, int c, int cl, int C,
                                              const T* stg) {
    const T tv = stg[(pair == 0 ? cl : C - 1 - cl) * (N + 1) + (N - 1 - h)];
    const T nv = (T)image_strip2(x, N, pair == 0 ? 1 : 2, h, c);
    return pair == 0 ? mk(tv, nv) : mk(nv, tv);
}

}  // namespace fftreg
}  // namespace adrt
