// Offset-axis FFTs of length L = 32 * L1 (L = 128 .. 1024) as a two-step (four-step) FFT held in registers.
//
//   X[k2 + 32 k1] = sum_h1 W_L^{h1 k2} W_L1^{h1 k1} sum_h2 x[h1 + L1 h2] W_32^{h2 k2}
//
// A block owns C columns of one (sinogram, pair) with L1 threads per column.  Step 1: thread (h1, c) loads the 32
// values x[h1 + L1 h2] of its column (fused with the quadrant packing / iota of the strips), does a 32-point DFT in
// registers and applies W_L^{h1 k2}.  One LDS exchange regroups the values by k2.  Step 2: each thread does 32 / L1
// DFTs of length L1 over h1 and writes X[k2 + 32 k1] (the inverse variant splits Re / Im into the quadrant strips and
// keeps rows h < N).  Global reads and writes move C consecutive columns per row.  Twiddles come from the plan's
// table W_L^m (m < L/2) staged in LDS.
#include "fftreg.hpp"

namespace adrt {

namespace {

using namespace fftreg;

// MODE 0: pack image strips (forward), 1: pack data strips (forward), 2: inverse + quadrant split
template <typename Tg, typename T, int L1, int C, int MODE>
__global__ __launch_bounds__(L1* C) void fft2_kernel(int n, int nsino, const void* src, void* dst, const cplx<T>* twg) {
    constexpr int L = 32 * L1;
    constexpr int P = C + 1;  // LDS pitch per row of the exchange
    constexpr bool INV = MODE == 2;
    constexpr int KPT = 32 / L1;  // k2 values per thread in step 2
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* S = reinterpret_cast<cplx<T>*>(smem_raw);  // L * P values
    cplx<T>* tw = S + L * P;                              // L/2 twiddles
    const int N = 1 << n;
    const int t = threadIdx.x;
    const int cl = t % C, h1 = t / C;
    // blocks of one sinogram / image on one XCD: an image is read once per quadrant of each pair, from that L2
    const int gx = N / C;
    const int lb = xcd_block(blockIdx.x, gx * 2 * nsino);
    if (lb < 0) return;
    const int c = (lb % gx) * C + cl;
    const int pair = (lb / gx) % 2, s = lb / (2 * gx);
    for (int i = t; i < L / 2; i += L1 * C) tw[i] = twg[i];
    const double* x = MODE == 0 ? reinterpret_cast<const double*>(src) + (size_t)s * N * N : nullptr;
    T* stg = reinterpret_cast<T*>(S);  // mode 0: the transposed strip's rows, before the exchange uses S
    if (MODE == 0) stage_rows<T, 16 * L1, C, L1 * C>(x, pair, c - cl, stg, t);  // N = 16 L1
    __syncthreads();
    cplx<T> w32[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) w32[m] = tw[m * L1];

    // ---- step 1: load x[h1 + L1 h2], 32-point DFT over h2, twiddle W_L^{h1 k2}
    cplx<T> v[32];
    if (MODE == 0) {
#pragma unroll
        for (int h2 = 0; h2 < 32; ++h2) {
            const int h = h1 + L1 * h2;
            // h < N (N = 16 L1); the zero padding rows h >= N contribute nothing
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
    // v[brev5(k2)] holds the sum for k2; apply W_L^{h1 k2} and store to S[(h1 + L1 k2) * P + cl]
#pragma unroll
    for (int k2 = 0; k2 < 32; ++k2) {
        cplx<T> y = v[brev_c<32>(k2)];
        if (k2 != 0 && h1 != 0) y = rot<T, INV>(tw_at(tw, h1 * k2, L), y);
        S[(h1 + L1 * k2) * P + cl] = y;
    }
    __syncthreads();

    // ---- step 2: thread (g, cl) owns k2 = g * KPT + u; L1-point DFT over h1
    const int g = h1;
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
        const int k2 = g * KPT + u;
        cplx<T> w[L1];
#pragma unroll
        for (int a = 0; a < L1; ++a) w[a] = S[(a + L1 * k2) * P + cl];
        dft_dif<T, L1, INV>(w, w32);
#pragma unroll
        for (int k1 = 0; k1 < L1; ++k1) {
            const cplx<T> X = w[brev_c<L1>(k1)];
            const int k = k2 + 32 * k1;
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

// ---- inverse FFT with the quadrant strips summed in pairs
//
// The combine adds, at image pixel (a, c), S0[N-1-c][a] + S1[N-1-a][c] + S2[a][c] + S3[N-1-c][N-1-a] (q0, q1 the real
// and imaginary parts of pair 0, q2, q3 of pair 1).  Two sums of that need no transpose between them:
//   U[h][a] = S0[h][a] + S3[h][N-1-a]   (pair 0 at column a, pair 1 at the mirrored column, same row)
//   V[a][c] = S1[N-1-a][c] + S2[a][c]   (pair 0 at the mirrored row, pair 1, same column)
// so the inverse writes the two planes U, V instead of the four strips and the combine reads U^T + V: half the strip
// traffic on both sides.  A block owns the column group g and its mirror gx-1-g of both pairs and runs four column
// FFTs (units) one after another, each with the thread -> (column, row) mapping chosen so that every pair of values
// that one sum adds sits in one thread: the mirror units take their columns reversed (thread cl at column N-1-c), and
// units 2 and 3 their rows mirrored within the step-2 split (k2 -> 31 - k2: row k -> N-1-k for the kept k < N).

// One unit: the inverse column FFT of Zp at column c as in fft2_kernel (MODE 2), the kept rows k < N scaled by 1 / L
// and rounded to Tg into re / im[u][k1], row k = k2 + 32 k1 with k2 = h1 * KPT + u, or 31 - (h1 * KPT + u) with rk.
template <typename Tg, typename T, int L1, int C>
__device__ __forceinline__ void inv_unit(const cplx<Tg>* Zp, int c, int h1, int cl, bool rk, cplx<T>* S,
                                         const cplx<T>* tw, const cplx<T> (&w32)[16], Tg (&re)[32 / L1][L1 / 2],
                                         Tg (&im)[32 / L1][L1 / 2]) {
    constexpr int L = 32 * L1, N = 16 * L1, P = C + 1, KPT = 32 / L1;
    cplx<T> v[32];
#pragma unroll
    for (int h2 = 0; h2 < 32; ++h2) v[h2] = cast<T>(Zp[(h1 + L1 * h2) * N + c]);
    dft_dif<T, 32, true>(v, w32);
#pragma unroll
    for (int k2 = 0; k2 < 32; ++k2) {
        cplx<T> y = v[brev_c<32>(k2)];
        if (k2 != 0 && h1 != 0) y = rot<T, true>(tw_at(tw, h1 * k2, L), y);
        S[(h1 + L1 * k2) * P + cl] = y;
    }
    __syncthreads();
    const T inv = (T)1 / (T)L;
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
        const int k2 = rk ? 31 - (h1 * KPT + u) : h1 * KPT + u;
        cplx<T> w[L1];
#pragma unroll
        for (int a = 0; a < L1; ++a) w[a] = S[(a + L1 * k2) * P + cl];
        dft_dif<T, L1, true>(w, w32);
#pragma unroll
        for (int k1 = 0; k1 < L1 / 2; ++k1) {  // k = k2 + 32 k1 < N = 16 L1
            const cplx<T> X = w[brev_c<L1>(k1)];
            re[u][k1] = (Tg)(X.x * inv);
            im[u][k1] = (Tg)(X.y * inv);
        }
    }
    __syncthreads();  // the next unit's step 1 overwrites S
}

// Z: (nsino, 2, L, N) packed pairs after the row pass.  UV: (nsino, 2, N, N), plane 0 = U, plane 1 = V.
template <typename Tg, typename T, int L1, int C>
__global__ __launch_bounds__(L1* C) void ifft_pair_kernel(int nsino, const void* src, void* dst, const cplx<T>* twg) {
    constexpr int L = 32 * L1, N = 16 * L1;
    constexpr int P = C + 1;
    constexpr int KPT = 32 / L1, H = L1 / 2;
    constexpr int HX = N / C / 2;  // column groups below N/2 (the launcher checks N / C is even)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cplx<T>* S = reinterpret_cast<cplx<T>*>(smem_raw);  // L * P values
    cplx<T>* tw = S + L * P;                              // L/2 twiddles
    const int t = threadIdx.x;
    const int cl = t % C, h1 = t / C;
    const int lb = xcd_block(blockIdx.x, HX * nsino);
    if (lb < 0) return;
    const int s = lb / HX;
    const int c = (lb % HX) * C + cl;  // units 0, 1
    const int m = N - 1 - c;           // units 2, 3: the mirrored group, columns reversed
    for (int i = t; i < L / 2; i += L1 * C) tw[i] = twg[i];
    __syncthreads();
    cplx<T> w32[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) w32[q] = tw[q * L1];
    const cplx<Tg>* Z0 = reinterpret_cast<const cplx<Tg>*>(src) + (size_t)s * 2 * L * N;
    Tg* U = reinterpret_cast<Tg*>(dst) + (size_t)s * 2 * N * N;
    Tg* V = U + N * N;
    Tg s0[KPT][H], keep[KPT][H], xr[KPT][H], xi[KPT][H];
#pragma unroll 1
    for (int unit = 0; unit < 4; ++unit) {
        // units 1 and 2 take the mirrored rows, units 2 and 3 the mirrored columns, units 1 and 3 pair 1
        const bool rk = unit == 1 || unit == 2;
        // the thread's indices as values the compiler must take as new in every unit: otherwise it computes the
        // addresses of all four units' loads and stores once before the loop and keeps them in registers
        int c_ = c, m_ = m, h1_ = h1;
        asm volatile("" : "+v"(c_), "+v"(m_), "+v"(h1_));
        inv_unit<Tg, T, L1, C>(Z0 + (unit & 1) * L * N, unit < 2 ? c_ : m_, h1_, cl, rk, S, tw, w32, xr, xi);
        if (unit == 0) {  // pair 0 at column c, rows k2 + 32 k1: S0 (kept for unit 3), S1
#pragma unroll
            for (int u = 0; u < KPT; ++u)
#pragma unroll
                for (int k1 = 0; k1 < H; ++k1) {
                    s0[u][k1] = xr[u][k1];
                    keep[u][k1] = xi[u][k1];
                }
        } else if (unit == 1) {  // pair 1 at column c, mirrored rows a: V[a][c] = S1[N-1-a][c] + S2[a][c]; keep S3
#pragma unroll
            for (int u = 0; u < KPT; ++u)
#pragma unroll
                for (int k1 = 0; k1 < H; ++k1) {
                    const int a = 31 - (h1_ * KPT + u) + 32 * k1;
                    V[a * N + c_] = keep[u][H - 1 - k1] + xr[u][k1];
                }
#pragma unroll
            for (int u = 0; u < KPT; ++u)
#pragma unroll
                for (int k1 = 0; k1 < H; ++k1) keep[u][k1] = xi[u][k1];
        } else if (unit == 2) {  // pair 0 at column m, mirrored rows h: U[h][m] = S0[h][m] + S3[h][c]; keep S1
#pragma unroll
            for (int u = 0; u < KPT; ++u)
#pragma unroll
                for (int k1 = 0; k1 < H; ++k1) {
                    const int h = 31 - (h1_ * KPT + u) + 32 * k1;
                    U[h * N + m_] = xr[u][k1] + keep[u][k1];
                    keep[u][k1] = xi[u][k1];
                }
        } else {  // pair 1 at column m, rows a: V[a][m] = S1[N-1-a][m] + S2[a][m], U[a][c] = S0[a][c] + S3[a][m]
#pragma unroll
            for (int u = 0; u < KPT; ++u)
#pragma unroll
                for (int k1 = 0; k1 < H; ++k1) {
                    const int a = h1_ * KPT + u + 32 * k1;
                    V[a * N + m_] = keep[u][H - 1 - k1] + xr[u][k1];
                    U[a * N + c_] = s0[u][k1] + xi[u][k1];
                }
        }
    }
}

template <typename Tg, typename T, int L1, int C, int MODE>
int launch_cfg(int n, const void* src, void* dst, int S, const void* tw, hipStream_t st) {
    constexpr int L = 32 * L1;
    const size_t lds = sizeof(cplx<T>) * ((size_t)L * (C + 1) + L / 2);
    const int N = 1 << n;
    ADRT_REQUIRE((long long)(N / C) * 2 * S <= 0x7fffffffLL, "fft2: too many blocks");
    hipLaunchKernelGGL((fft2_kernel<Tg, T, L1, C, MODE>), dim3(xcd_grid((long long)(N / C) * 2 * S)), dim3(L1 * C), lds,
                       st, n, S, src, dst, (const cplx<T>*)tw);
    return launch_check("fft2_kernel");
}

template <typename Tg, typename T, int L1, int C>
int launch_pair_cfg(int n, const void* src, void* dst, int S, const void* tw, hipStream_t st) {
    constexpr int L = 32 * L1;
    const size_t lds = sizeof(cplx<T>) * ((size_t)L * (C + 1) + L / 2);
    const int N = 1 << n;
    ADRT_REQUIRE(N == 16 * L1 && N % (2 * C) == 0, "ifft_pair: needs N = 16 L1 and an even number of column groups");
    ADRT_REQUIRE((long long)(N / C / 2) * S <= 0x7fffffffLL, "ifft_pair: too many blocks");
    hipLaunchKernelGGL((ifft_pair_kernel<Tg, T, L1, C>), dim3(xcd_grid((long long)(N / C / 2) * S)), dim3(L1 * C), lds,
                       st, S, src, dst, (const cplx<T>*)tw);
    return launch_check("ifft_pair_kernel");
}

// columns per block of fft2_kernel for n and the compute type (the launch_mode table below)
constexpr int fft2_cols(int n, bool f64) {
    return n == 6 ? (f64 ? 32 : 64) : n == 7 ? (f64 ? 16 : 32) : n == 8 ? (f64 ? 8 : 16) : n == 9 ? (f64 ? 4 : 16) : 0;
}

template <typename Tg, typename T>
int launch_pair(int n, const void* src, void* dst, int S, const void* tw, hipStream_t st) {
    constexpr bool F64 = sizeof(T) == 8;
    switch (n) {
        case 7: return launch_pair_cfg<Tg, T, 8, fft2_cols(7, F64)>(n, src, dst, S, tw, st);
        case 8: return launch_pair_cfg<Tg, T, 16, fft2_cols(8, F64)>(n, src, dst, S, tw, st);
        case 9: return launch_pair_cfg<Tg, T, 32, fft2_cols(9, F64)>(n, src, dst, S, tw, st);
        default: set_error("ifft_pair: n must be 7..9"); return E_INVALID;
    }
}

template <typename Tg, typename T, int MODE>
int launch_mode(int n, const void* src, void* dst, int S, const void* tw, hipStream_t st) {
    constexpr bool F64 = sizeof(T) == 8;
    switch (n) {
        case 6: return launch_cfg<Tg, T, 4, fft2_cols(6, F64), MODE>(n, src, dst, S, tw, st);
        case 7: return launch_cfg<Tg, T, 8, fft2_cols(7, F64), MODE>(n, src, dst, S, tw, st);
        case 8: return launch_cfg<Tg, T, 16, fft2_cols(8, F64), MODE>(n, src, dst, S, tw, st);
        // float at N = 512: 16 columns (128-byte row segments) in one 143 KiB block per CU measured 11% faster for the
        // FFTs and combine of a pass than 8 columns in two 78 KiB blocks
        case 9: return launch_cfg<Tg, T, 32, fft2_cols(9, F64), MODE>(n, src, dst, S, tw, st);
        default: set_error("fft2: n must be 6..9"); return E_INVALID;
    }
}

}  // namespace

bool fft2_supported(int n) { return n >= 6 && n <= 9; }

int ifft_pair_blocks(int n, int prec) {
    if (prec != PREC_F32 || !fft2_supported(n)) return 0;
    const int groups = (1 << n) / fft2_cols(n, false);
    return groups % 2 == 0 ? groups / 2 : 0;
}

int launch_ifft_pair(int n, const void* Z, void* UV, int S, int prec, const void* tw, hipStream_t st) {
    if (S <= 0) return OK;
    ADRT_REQUIRE(ifft_pair_blocks(n, prec) > 0, "ifft_pair: float, 7 <= n <= 9 only");
    return launch_pair<float, float>(n, Z, UV, S, tw, st);
}

int launch_fft2(int n, int mode, const void* src, void* dst, int S, int prec, const void* tw, hipStream_t st) {
    if (S <= 0) return OK;
    ADRT_REQUIRE(fft2_supported(n), "fft2: n must be 6..9");
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
