// Offset-axis FFTs of length L = 32 * L1 with L1 = 8 * BN (L = 2048, 4096: N = 1024, 2048) as a three-step FFT held
// in registers.
//
//   X[k2 + 32 k1] = sum_h1 W_L1^{h1 k1} Y1[h1][k2],      Y1[h1][k2] = W_L^{h1 k2} sum_h2 x[h1 + L1 h2] W_32^{h2 k2}
//   and, with h1 = a + 8 b, k1 = kb + BN ka:
//   sum_h1 W_L1^{h1 k1} Y1[h1][k2] = sum_a W_8^{a ka} W_L1^{a kb} sum_b W_BN^{b kb} Y1[a + 8 b][k2]
//
// A block owns C columns of one (sinogram, pair) with L1 threads per column; every thread holds 32 values in each
// step.  Step 1: thread (h1, c) loads x[h1 + L1 h2] (fused with the quadrant packing / iota of the strips, as the
// two-step kernel of fft2.hip), does a 32-point DFT and applies W_L^{h1 k2}; the column goes to LDS.  Step 2: each
// thread takes 256 / L1 of the (k2, a) groups, does their BN-point DFTs over b and applies W_L1^{a kb}; the column goes
// back to LDS regrouped by (k2, kb).  Step 3: each thread takes 4 of the (k2, kb) groups, does their 8-point DFTs
// over a and writes X[k2 + 32 (kb + BN ka)] to HBM (the inverse splits Re / Im into the quadrant strips and keeps
// rows h < N).  The whole column sits in LDS (L values), so a block holds C = 128 KiB / (L * value size) columns:
// 8 / 4 for float and 4 / 2 for double at N = 1024 / 2048.  Twiddles W_L^m come from the plan's table in HBM.
#include "fftreg.hpp"

namespace adrt {

namespace {

using namespace fftreg;

template <typename Tg, typename T, int L1, int C, int MODE>
__global__ __launch_bounds__(L1* C) void fft3_kernel(int n, int nsino, const void* src, void* dst,
                                                     const cplx<T>* twg) {
    constexpr int L = 32 * L1;
    constexpr int BN = L1 / 8;
    constexpr int NT2 = 256 / L1;  // (k2, a) groups per thread in step 2
    constexpr bool INV = MODE == 2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* S = reinterpret_cast<cplx<T>*>(smem_raw);  // L * C values: entry (i, k2) of column cl at (i + L1 k2) C + cl
    const int N = 1 << n;
    const int t = threadIdx.x;
    const int cl = t % C, h1 = t / C;
    const int gx = N / C;
    const int lb = xcd_block(blockIdx.x, gx * 2 * nsino);
    if (lb < 0) return;
    const int c = (lb % gx) * C + cl;
    const int pair = (lb / gx) % 2, s = lb / (2 * gx);
    cplx<T> w32[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) w32[m] = twg[m * L1];

    // ---- step 1: x[h1 + L1 h2] -> 32-point DFT over h2 -> W_L^{h1 k2} -> LDS (h1, k2)
    cplx<T> v[32];
    if (MODE == 0) {
        const double* x = reinterpret_cast<const double*>(src) + (size_t)s * N * N;
        T* stg = reinterpret_cast<T*>(S);  // the transposed strip's rows, before step 1 uses S
        stage_rows<T, 16 * L1, C, L1 * C>(x, pair, c - cl, stg, t);  // N = 16 L1
        __syncthreads();
#pragma unroll
        for (int h2 = 0; h2 < 32; ++h2) {
            const int h = h1 + L1 * h2;
            // h < N; the zero padding rows h >= N contribute nothing
            v[h2] = h2 < 16 ? strip_pair<T>(x, N, pair, h, c, cl, C, stg) : mk((T)0, (T)0);
        }
        __syncthreads();  // every thread has its column before step 1 overwrites the staged rows
    } else if (MODE == 1) {
        const size_t qsz = (size_t)(2 * N - 1) * N;
        const double* ba = reinterpret_cast<const double*>(src) + ((size_t)s * 4 + 2 * pair) * qsz;
        const double* bb = ba + qsz;
#pragma unroll
        for (int h2 = 0; h2 < 32; ++h2) {
            const int h = h1 + L1 * h2;
            if (h == N) {
                v[h2] = mk((T)0, (T)0);
            } else {
                const int r = h < N ? h + N - 1 : h - N - 1;  // sinogram row of torus row h
                v[h2] = mk((T)ba[(size_t)r * N + c], (T)bb[(size_t)r * N + c]);
            }
        }
    } else {
        const cplx<Tg>* Z = reinterpret_cast<const cplx<Tg>*>(src) + ((size_t)s * 2 + pair) * L * N;
#pragma unroll
        for (int h2 = 0; h2 < 32; ++h2) v[h2] = cast<T>(Z[(size_t)(h1 + L1 * h2) * N + c]);
    }
    dft_dif<T, 32, INV>(v, w32);
#pragma unroll
    for (int k2 = 0; k2 < 32; ++k2) {
        cplx<T> y = v[brev_c<32>(k2)];
        if (k2 != 0 && h1 != 0) y = rot<T, INV>(tw_at(twg, h1 * k2, L), y);
        S[(h1 + L1 * k2) * C + cl] = y;
    }
    __syncthreads();

    // ---- step 2: group (k2, a), q = h1 + L1 u: BN-point DFT over b of Y1[a + 8 b][k2], times W_L1^{a kb}
    cplx<T> z[NT2][BN];
#pragma unroll
    for (int u = 0; u < NT2; ++u) {
        const int q = h1 + L1 * u;
        const int a = q % 8, k2 = q / 8;
#pragma unroll
        for (int b = 0; b < BN; ++b) z[u][b] = S[(a + 8 * b + L1 * k2) * C + cl];
        dft_dif<T, BN, INV>(z[u], w32);
    }
    __syncthreads();  // every group has read its column entries; the regrouped values overwrite them
#pragma unroll
    for (int u = 0; u < NT2; ++u) {
        const int q = h1 + L1 * u;
        const int a = q % 8, k2 = q / 8;
#pragma unroll
        for (int kb = 0; kb < BN; ++kb) {
            cplx<T> y = z[u][brev_c<BN>(kb)];
            if (a != 0 && kb != 0) y = rot<T, INV>(tw_at(twg, 32 * a * kb, L), y);
            S[(kb + BN * a + L1 * k2) * C + cl] = y;  // regrouped: (k2, kb) groups contiguous in a
        }
    }
    __syncthreads();

    // ---- step 3: group (k2, kb), p = h1 + L1 u: 8-point DFT over a -> X[k2 + 32 (kb + BN ka)]
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int p = h1 + L1 * u;
        const int kb = p % BN, k2 = p / BN;
        cplx<T> w[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) w[a] = S[(kb + BN * a + L1 * k2) * C + cl];
        dft_dif<T, 8, INV>(w, w32);
#pragma unroll
        for (int ka = 0; ka < 8; ++ka) {
            const cplx<T> X = w[brev_c<8>(ka)];
            const int k = k2 + 32 * (kb + BN * ka);
            if (MODE != 2) {
                cplx<Tg>* Y = reinterpret_cast<cplx<Tg>*>(dst) + ((size_t)s * 2 + pair) * L * N;
                Y[(size_t)k * N + c] = cast<Tg>(X);
            } else if (k < N) {
                const T inv = (T)1 / (T)L;
                Tg* Sa = reinterpret_cast<Tg*>(dst) + ((size_t)s * 4 + 2 * pair) * N * N;
                Tg* Sb = Sa + (size_t)N * N;
                Sa[(size_t)k * N + c] = (Tg)(X.x * inv);
                Sb[(size_t)k * N + c] = (Tg)(X.y * inv);
            }
        }
    }
}

template <typename Tg, typename T, int L1, int C, int MODE>
int launch_cfg(int n, const void* src, void* dst, int S, const void* tw, hipStream_t st) {
    constexpr int L = 32 * L1;
    constexpr size_t lds = sizeof(cplx<T>) * (size_t)L * C;
    static_assert(lds <= 128 * 1024, "fft3: column buffer exceeds 128 KiB of LDS");
    const int N = 1 << n;
    const long long nblk = (long long)(N / C) * 2 * S;
    ADRT_REQUIRE(nblk <= 0x7fffffffLL, "fft3: too many blocks");
    hipLaunchKernelGGL((fft3_kernel<Tg, T, L1, C, MODE>), dim3(xcd_grid(nblk)), dim3(L1 * C), lds, st, n, S, src, dst,
                       (const cplx<T>*)tw);
    return launch_check("fft3_kernel");
}

template <typename Tg, typename T, int MODE>
int launch_mode(int n, const void* src, void* dst, int S, const void* tw, hipStream_t st) {
    constexpr bool F64 = sizeof(T) == 8;
    switch (n) {
        case 10: return launch_cfg<Tg, T, 64, (F64 ? 4 : 8), MODE>(n, src, dst, S, tw, st);
        case 11: return launch_cfg<Tg, T, 128, (F64 ? 2 : 4), MODE>(n, src, dst, S, tw, st);
        default: set_error("fft3: n must be 10 or 11"); return E_INVALID;
    }
}

}  // namespace

bool fft3_supported(int n) { return n == 10 || n == 11; }

int launch_fft3(int n, int mode, const void* src, void* dst, int S, int prec, const void* tw, hipStream_t st) {
    if (S <= 0) return OK;
    ADRT_REQUIRE(fft3_supported(n), "fft3: n must be 10 or 11");
    if (prec == PREC_F64) {
        if (mode == 0) return launch_mode<double, double, 0>(n, src, dst, S, tw, st);
        if (mode == 1) return launch_mode<double, double, 1>(n, src, dst, S, tw, st);
        return launch_mode<double, double, 2>(n, src, dst, S, tw, st);
    }
    if (mode == 0) return launch_mode<float, float, 0>(n, src, dst, S, tw, st);
    if (mode == 1) return launch_mode<float, float, 1>(n, src, dst, S, tw, st);
    return launch_mode<float, float, 2>(n, src, dst, S, tw, st);
}

}  // namespace adrt
