// Offset-axis FFTs of the packed quadrant strips, fused with the quadrant packing (forward) and the quadrant
// split (inverse).  One block owns C columns of one (sinogram, pair); the L = 2N offsets of each column sit in LDS
// and are transformed by an in-place radix-2 decimation-in-time FFT whose input is loaded in bit-reversed order,
// so the output comes out natural.  Twiddles W^m = exp(-2 pi i m / L) come from a table in HBM (L2 resident).
#include "passes.hpp"

namespace adrt {

namespace {

constexpr int FFT_THREADS = 512;
constexpr size_t FFT_LDS_BUDGET = 128 * 1024;  // of the 160 KiB of LDS per CU

__device__ __forceinline__ int brev_bits(int v, int bits) { return bits ? (int)(__brev((unsigned)v) >> (32 - bits)) : 0; }

template <typename T>
__device__ __forceinline__ void fft_stages(cplx<T>* buf, int C, int stride, int logL, const cplx<T>* tw, bool inverse) {
    const int L = 1 << logL;
    const int halfL = L >> 1;
    for (int s = 1; s <= logL; ++s) {
        const int half = 1 << (s - 1);
        for (int b = threadIdx.x; b < C * halfL; b += FFT_THREADS) {
            const int col = b >> (logL - 1);
            const int j = b & (halfL - 1);
            const int jj = j & (half - 1);
            const int i0 = ((j >> (s - 1)) << s) | jj;
            const int i1 = i0 + half;
            cplx<T> w = tw[jj << (logL - s)];
            if (inverse) w.y = -w.y;
            cplx<T>* X = buf + (size_t)col * stride;
            const cplx<T> u = X[i0];
            const cplx<T> v = mul(w, X[i1]);
            X[i0] = add(u, v);
            X[i1] = sub(u, v);
        }
        __syncthreads();
    }
}

// image strips: value of quadrant q at (h, col), h < N (adrtcore.py:29-39 through fastspec.Tq)
__device__ __forceinline__ double image_strip(const double* x, int N, int q, int h, int col) {
    switch (q) {
        case 0: return x[(size_t)col * N + (N - 1 - h)];
        case 1: return x[(size_t)(N - 1 - h) * N + col];
        case 2: return x[(size_t)h * N + col];
        default: return x[(size_t)(N - 1 - col) * N + (N - 1 - h)];
    }
}

template <typename Tg, typename T>
__global__ __launch_bounds__(FFT_THREADS) void pack_fft_kernel(int n, int C, int mode, const double* src, void* Yv,
                                                                const cplx<T>* tw) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
    const int N = 1 << n, logL = n + 1, L = 2 * N;
    const int stride = C == 1 ? L : L + 1;
    const int c0 = blockIdx.x * C;
    const int pair = blockIdx.y;
    const int s = blockIdx.z;
    const int qa = 2 * pair, qb = 2 * pair + 1;
    const int tot = C * L;
    if (mode == 0) {
        const double* x = src + (size_t)s * N * N;
        // quadrants 1 and 2 read image rows (contiguous in col), 0 and 3 read image columns (contiguous in h)
        for (int part = 0; part < 2; ++part) {
            const int q = part == 0 ? qa : qb;
            const bool h_fast = (q == 0 || q == 3);
            for (int idx = threadIdx.x; idx < tot; idx += FFT_THREADS) {
                int h, cl;
                if (h_fast) {
                    h = idx % L;
                    cl = idx / L;
                } else {
                    cl = idx % C;
                    h = idx / C;
                }
                const T v = h < N ? (T)image_strip(x, N, q, h, c0 + cl) : (T)0;
                cplx<T>& e = buf[(size_t)cl * stride + brev_bits(h, logL)];
                if (part == 0)
                    e.x = v;
                else
                    e.y = v;
            }
        }
    } else {
        const size_t qsz = (size_t)(2 * N - 1) * N;
        const double* ba = src + ((size_t)s * 4 + qa) * qsz;
        const double* bb = src + ((size_t)s * 4 + qb) * qsz;
        for (int idx = threadIdx.x; idx < tot; idx += FFT_THREADS) {
            const int cl = idx % C;
            const int h = idx / C;
            T va = 0, vb = 0;
            if (h != N) {
                const int r = h < N ? h + N - 1 : h - N - 1;  // sinogram row of torus row h
                va = (T)ba[(size_t)r * N + c0 + cl];
                vb = (T)bb[(size_t)r * N + c0 + cl];
            }
            buf[(size_t)cl * stride + brev_bits(h, logL)] = mk(va, vb);
        }
    }
    __syncthreads();
    fft_stages(buf, C, stride, logL, tw, false);
    cplx<Tg>* Y = reinterpret_cast<cplx<Tg>*>(Yv) + ((size_t)s * 2 + pair) * L * N;
    for (int idx = threadIdx.x; idx < tot; idx += FFT_THREADS) {
        const int cl = idx % C;
        const int k = idx / C;
        Y[(size_t)k * N + c0 + cl] = cast<Tg>(buf[(size_t)cl * stride + k]);
    }
}

template <typename Tg, typename T>
__global__ __launch_bounds__(FFT_THREADS) void ifft_unpack_kernel(int n, int C, const void* Zv, Tg* Sq,
                                                                   const cplx<T>* tw) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
    const int N = 1 << n, logL = n + 1, L = 2 * N;
    const int stride = C == 1 ? L : L + 1;
    const int c0 = blockIdx.x * C;
    const int pair = blockIdx.y;
    const int s = blockIdx.z;
    const int tot = C * L;
    const cplx<Tg>* Z = reinterpret_cast<const cplx<Tg>*>(Zv) + ((size_t)s * 2 + pair) * L * N;
    for (int idx = threadIdx.x; idx < tot; idx += FFT_THREADS) {
        const int cl = idx % C;
        const int k = idx / C;
        buf[(size_t)cl * stride + brev_bits(k, logL)] = cast<T>(Z[(size_t)k * N + c0 + cl]);
    }
    __syncthreads();
    fft_stages(buf, C, stride, logL, tw, true);
    const T inv = (T)1 / (T)L;
    Tg* Sa = Sq + ((size_t)s * 4 + 2 * pair) * N * N;
    Tg* Sb = Sa + (size_t)N * N;
    for (int idx = threadIdx.x; idx < C * N; idx += FFT_THREADS) {
        const int cl = idx % C;
        const int h = idx / C;
        const cplx<T> v = buf[(size_t)cl * stride + h];
        Sa[(size_t)h * N + c0 + cl] = (Tg)(v.x * inv);
        Sb[(size_t)h * N + c0 + cl] = (Tg)(v.y * inv);
    }
}

// out[a][c] = (S0[N-1-c][a] + S1[N-1-a][c]) + (S2[a][c] + S3[N-1-c][N-1-a]), optionally y - out.
template <typename Tg, typename To>
__global__ __launch_bounds__(256) void combine_kernel(int n, const Tg* Sq, const double* y, To* out) {
    __shared__ double t0[32][33], t3[32][33];
    const int N = 1 << n;
    const int s = blockIdx.z;
    const int a0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const Tg* S0 = Sq + (size_t)s * 4 * N * N;
    const Tg* S1 = S0 + (size_t)N * N;
    const Tg* S2 = S1 + (size_t)N * N;
    const Tg* S3 = S2 + (size_t)N * N;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // the untransposed strips and y first, so that their loads are in flight across the tile's barrier
    double v1[4], v2[4], yv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = a0 + ty + 8 * i, c = c0 + tx;
        const bool in = a < N && c < N;
        v1[i] = in ? (double)S1[(size_t)(N - 1 - a) * N + c] : 0.0;
        v2[i] = in ? (double)S2[(size_t)a * N + c] : 0.0;
        yv[i] = (in && y) ? y[((size_t)s * N + a) * N + c] : 0.0;
    }
    // t0[c][a] = S0[N-1-(c0+c)][a0+a], t3[c][a] = S3[N-1-(c0+c)][N-1-(a0+a)]: contiguous along a
    for (int r = ty; r < 32; r += 8) {
        if (c0 + r < N && a0 + tx < N) {
            const size_t row = (size_t)(N - 1 - (c0 + r)) * N;
            t0[r][tx] = (double)S0[row + a0 + tx];
            t3[r][tx] = (double)S3[row + (N - 1 - (a0 + tx))];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = ty + 8 * i, a = a0 + r, c = c0 + tx;
        if (a >= N || c >= N) continue;
        const double v = (t0[tx][r] + v1[i]) + (v2[i] + t3[tx][r]);
        const size_t o = ((size_t)s * N + a) * N + c;
        out[o] = (To)(y ? yv[i] - v : v);
    }
}

// The combine of the paired planes (launch_ifft_pair): out[a][c] = U[N-1-c][a] + V[a][c] (or y - that), U read
// through a transposed LDS tile.
template <typename Tg, typename To>
__global__ __launch_bounds__(256) void combine_pair_kernel(int n, const Tg* UV, const double* y, To* out) {
    __shared__ double tu[32][33];
    const int N = 1 << n;
    const int s = blockIdx.z;
    const int a0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const Tg* U = UV + (size_t)s * 2 * N * N;
    const Tg* V = U + (size_t)N * N;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // V and y first, so that their loads are in flight across the tile's barrier
    double vv[4], yv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int a = a0 + ty + 8 * i, c = c0 + tx;
        const bool in = a < N && c < N;
        vv[i] = in ? (double)V[(size_t)a * N + c] : 0.0;
        yv[i] = (in && y) ? y[((size_t)s * N + a) * N + c] : 0.0;
    }
    // tu[c][a] = U[N-1-(c0+c)][a0+a]: contiguous along a
    for (int r = ty; r < 32; r += 8)
        if (c0 + r < N && a0 + tx < N) tu[r][tx] = (double)U[(size_t)(N - 1 - (c0 + r)) * N + a0 + tx];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = ty + 8 * i, a = a0 + r, c = c0 + tx;
        if (a >= N || c >= N) continue;
        const double v = tu[tx][r] + vv[i];
        const size_t o = ((size_t)s * N + a) * N + c;
        out[o] = (To)(y ? yv[i] - v : v);
    }
}

int fft_cols(int n, size_t elem) {
    const size_t L = (size_t)2 << n;
    const size_t N = (size_t)1 << n;
    size_t C = FFT_LDS_BUDGET / ((L + 1) * elem);
    if (C < 1) C = 1;
    if (C > N) C = N;
    if (C > 64) C = 64;
    size_t p = 1;
    while (p * 2 <= C) p *= 2;  // power of two so that it divides N
    return (int)p;
}

template <typename Tg, typename T>
int pack_impl(int n, int mode, const double* src, void* Y, int S, const void* tw, hipStream_t st) {
    const int C = fft_cols(n, sizeof(cplx<T>));
    const int N = 1 << n, L = 2 * N;
    const size_t lds = sizeof(cplx<T>) * (size_t)C * (C == 1 ? L : L + 1);
    hipLaunchKernelGGL((pack_fft_kernel<Tg, T>), dim3(N / C, 2, S), dim3(FFT_THREADS), lds, st, n, C, mode, src, Y,
                       (const cplx<T>*)tw);
    return launch_check("pack_fft_kernel");
}

template <typename Tg, typename T>
int ifft_impl(int n, const void* Z, void* Sq, int S, const void* tw, hipStream_t st) {
    const int C = fft_cols(n, sizeof(cplx<T>));
    const int N = 1 << n, L = 2 * N;
    const size_t lds = sizeof(cplx<T>) * (size_t)C * (C == 1 ? L : L + 1);
    hipLaunchKernelGGL((ifft_unpack_kernel<Tg, T>), dim3(N / C, 2, S), dim3(FFT_THREADS), lds, st, n, C, Z, (Tg*)Sq,
                       (const cplx<T>*)tw);
    return launch_check("ifft_unpack_kernel");
}

}  // namespace

size_t twiddle_bytes(int n, int prec) {
    const size_t half = (size_t)1 << n;  // L / 2 = N
    return half * (prec == PREC_F64 ? sizeof(c128) : sizeof(c64));
}

int launch_pack_fft(int n, int mode, const double* src, void* Y, int S, int prec, const void* tw, hipStream_t st) {
    ADRT_REQUIRE(n >= 1 && n <= 11, "pack_fft: need 1 <= n <= 11");
    ADRT_REQUIRE(mode == 0 || mode == 1, "pack_fft: mode must be 0 (image) or 1 (data)");
    if (S <= 0) return OK;
    if (fft2_supported(n)) return launch_fft2(n, mode, src, Y, S, prec, tw, st);
    if (fft3_supported(n)) return launch_fft3(n, mode, src, Y, S, prec, tw, st);
    if (prec == PREC_F64) return pack_impl<double, double>(n, mode, src, Y, S, tw, st);
    return pack_impl<float, float>(n, mode, src, Y, S, tw, st);
}

int launch_ifft_unpack(int n, const void* Z, void* Sq, int S, int prec, const void* tw, hipStream_t st) {
    ADRT_REQUIRE(n >= 1 && n <= 11, "ifft_unpack: need 1 <= n <= 11");
    if (S <= 0) return OK;
    if (fft2_supported(n)) return launch_fft2(n, 2, Z, Sq, S, prec, tw, st);
    if (fft3_supported(n)) return launch_fft3(n, 2, Z, Sq, S, prec, tw, st);
    if (prec == PREC_F64) return ifft_impl<double, double>(n, Z, Sq, S, tw, st);
    return ifft_impl<float, float>(n, Z, Sq, S, tw, st);
}

int launch_combine(int n, const void* Sq, int prec, const double* y, void* out, int S, hipStream_t st, bool out_f32) {
    ADRT_REQUIRE(n >= 1 && n <= 11, "combine: need 1 <= n <= 11");
    if (S <= 0) return OK;
    const int N = 1 << n;
    const int T = N < 32 ? 1 : N / 32;
    const dim3 grid(T, T, S);
    if (prec == PREC_F64 && !out_f32)
        hipLaunchKernelGGL((combine_kernel<double, double>), grid, dim3(256), 0, st, n, (const double*)Sq, y,
                           (double*)out);
    else if (prec == PREC_F64)
        hipLaunchKernelGGL((combine_kernel<double, float>), grid, dim3(256), 0, st, n, (const double*)Sq, y,
                           (float*)out);
    else if (!out_f32)
        hipLaunchKernelGGL((combine_kernel<float, double>), grid, dim3(256), 0, st, n, (const float*)Sq, y,
                           (double*)out);
    else
        hipLaunchKernelGGL((combine_kernel<float, float>), grid, dim3(256), 0, st, n, (const float*)Sq, y,
                           (float*)out);
    return launch_check("combine_kernel");
}

int launch_combine_pair(int n, const void* UV, int prec, const double* y, void* out, int S, hipStream_t st,
                        bool out_f32) {
    ADRT_REQUIRE(n >= 5 && n <= 11, "combine_pair: need 5 <= n <= 11");
    if (S <= 0) return OK;
    const int T = (1 << n) / 32;
    const dim3 grid(T, T, S);
    if (prec == PREC_F64 && !out_f32)
        hipLaunchKernelGGL((combine_pair_kernel<double, double>), grid, dim3(256), 0, st, n, (const double*)UV, y,
                           (double*)out);
    else if (prec == PREC_F64)
        hipLaunchKernelGGL((combine_pair_kernel<double, float>), grid, dim3(256), 0, st, n, (const double*)UV, y,
                           (float*)out);
    else if (!out_f32)
        hipLaunchKernelGGL((combine_pair_kernel<float, double>), grid, dim3(256), 0, st, n, (const float*)UV, y,
                           (double*)out);
    else
        hipLaunchKernelGGL((combine_pair_kernel<float, float>), grid, dim3(256), 0, st, n, (const float*)UV, y,
                           (float*)out);
    return launch_check("combine_pair_kernel");
}

}  // namespace adrt
