// Image-space kernels of the outer iteration (fastops.py:247-258) and the level-1 preconditioners
// (fasthybrid.py:81-130): Chebyshev updates, the G_1 stencil step, the lorr projector.  All in double.
#include "passes.hpp"

namespace adrt {

namespace {

constexpr int IMG_THREADS = 256;

inline unsigned grid_for(long long count) {
    long long g = (count + IMG_THREADS - 1) / IMG_THREADS;
    if (g > (1LL << 20)) g = 1LL << 20;
    return (unsigned)(g < 1 ? 1 : g);
}

__global__ void scale_kernel(const double* x, double alpha, double* out, long long count) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        out[i] = alpha * x[i];
}

__global__ void cheb_update_kernel(const double* x, const double* xp, const double* r, double a, double w1, double w0,
                                   double* out, long long count) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
        out[i] = x[i] + w1 * (a * r[i]) + w0 * (x[i] - (xp ? xp[i] : 0.0));
}

// (G_1 x)[i][j] = 8 x[i][j] + ((2 x[i][j^1] + x[i-1][j^1]) + x[i+1][j^1]) + ((2 x[i^1][j] + x[i^1][j-1]) + x[i^1][j+1])
// with zeros outside the image: 8x + J(K x) + (J(K x^T))^T of fasthybrid.py:81-98, summed in that order.
__device__ __forceinline__ double g1_at(const double* x, int N, int i, int j) {
    const int jp = j ^ 1, ip = i ^ 1;
    double t1 = 2.0 * x[(size_t)i * N + jp];
    if (i > 0) t1 += x[(size_t)(i - 1) * N + jp];
    if (i < N - 1) t1 += x[(size_t)(i + 1) * N + jp];
    double t2 = 2.0 * x[(size_t)ip * N + j];
    if (j > 0) t2 += x[(size_t)ip * N + j - 1];
    if (j < N - 1) t2 += x[(size_t)ip * N + j + 1];
    return 8.0 * x[(size_t)i * N + j] + t1 + t2;
}

__global__ void g1_step_kernel(int n, const double* y, const double* x, const double* xp, double a, double w1,
                               double w0, double* out, long long count) {
    const int N = 1 << n;
    const long long NN = (long long)N * N;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count; e += (long long)gridDim.x * blockDim.x) {
        const long long s = e / NN;
        const int i = (int)((e % NN) / N), j = (int)(e % N);
        const double* xs = x + s * NN;
        const double r = y[e] - g1_at(xs, N, i, j);
        const double dx = xp ? (x[e] - xp[e]) : (x[e] - 0.0);
        out[e] = x[e] + w1 * (a * r) + w0 * dx;
    }
}

// (y - Pi y)/4 + (Pi y)/12 with Pi the 2 x 2 cell mean (reshape(N/2,2,N/2,2).mean over the cell)
__global__ void lorr_kernel(int n, const double* y, double* out, long long count) {
    const int N = 1 << n;
    const long long NN = (long long)N * N;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count; e += (long long)gridDim.x * blockDim.x) {
        const long long s = e / NN;
        const int i = (int)((e % NN) / N), j = (int)(e % N);
        const double* ys = y + s * NN;
        const int i0 = i & ~1, j0 = j & ~1;
        const double cell = ((ys[(size_t)i0 * N + j0] + ys[(size_t)i0 * N + j0 + 1]) +
                             (ys[(size_t)(i0 + 1) * N + j0] + ys[(size_t)(i0 + 1) * N + j0 + 1])) / 4.0;
        out[e] = (y[e] - cell) / 4.0 + cell / 12.0;
    }
}

// g1_inv fused into one launch.  A block owns a TILE x TILE output tile and iterates the Chebyshev solve of G_1 on the
// tile grown by a halo of HALO >= deg - 1 pixels (even, so the region starts on an even pixel and 2 x 2 cells are
// aligned with G_1's pair structure): the stencil reaches one pixel per step, so after step k only the tile dilated
// by (deg - 1 - k) must be exact, and the cells outside that shrinking square are skipped.
//
// Each thread owns 2 x 2 cells (up to BPT of them) and keeps their y, current and previous iterate in registers.  For
// a cell (i, j), G_1 x = 8 x + [1 2 1] down the partner column j^1 (rows i-1..i+1) + [1 2 1] along the partner row i^1
// (columns j-1..j+1) (fasthybrid.py:81-98); inside a 2 x 2 cell both partners are in the cell, so a step reads only
// eight neighbours from LDS (rows above / below, columns left / right of the cell) instead of seven per pixel.  The
// iterate is double-buffered in LDS (step k reads buffer (k-1)&1 and writes k&1), one barrier per step.  Pixels
// outside the image and a one-pixel border around the region hold zero, the stencil's zero extension.  The epilogue
// optionally scales (mode 1) or applies the outer Chebyshev update (mode 2), so P and the outer step cost one pass.
constexpr int G1_MAXH = 16;
constexpr int G1_THREADS = 512;

struct G1Args {
    int n, deg, mode;
    const double* y;   // right-hand side (S, N, N)
    const double* x;   // outer iterate (mode 2)
    const double* xp;  // outer previous iterate (mode 2, may be null = 0)
    double a_out, w1, w0;
    double* out;
    double om[32];     // inner omega_k, k = 1..deg-1
};

template <int TILE, int HALO>
struct G1Cfg {
    static constexpr int R = TILE + 2 * HALO;  // region side (even)
    static constexpr int RB = R / 2;           // 2 x 2 cells per region side
    static constexpr int RP = R + 2;           // LDS pitch: region plus a zero border on each side
    static constexpr int BUF = RP * RP;
    static constexpr int NB = RB * RB;
    static constexpr int BPT = (NB + G1_THREADS - 1) / G1_THREADS;
    static constexpr size_t LDS = 2 * BUF * sizeof(double);
};

template <int TILE, int HALO>
__global__ __launch_bounds__(G1_THREADS) void g1_fused_kernel(G1Args A) {
    using C = G1Cfg<TILE, HALO>;
    constexpr int RB = C::RB, RP = C::RP, BUF = C::BUF, NB = C::NB, BPT = C::BPT;
    extern __shared__ double X[];  // two buffers of BUF doubles
    const int N = 1 << A.n;
    const int H = A.deg - 1;  // <= HALO
    const int s = blockIdx.z;
    const int gi0 = blockIdx.y * TILE - HALO, gj0 = blockIdx.x * TILE - HALO;
    const double* ys = A.y + (size_t)s * N * N;
    const double a = 2.0 / 20.0;
    for (int i = threadIdx.x; i < 2 * BUF; i += G1_THREADS) X[i] = 0.0;
    __syncthreads();
    double yv[BPT][4], xc[BPT][4], xq[BPT][4];
    unsigned inm[BPT];
#pragma unroll
    for (int e = 0; e < BPT; ++e) {
        const int b = threadIdx.x + G1_THREADS * e;
        const int bi = b / RB, bj = b % RB;
        inm[e] = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int ri = 2 * bi + (c >> 1), rj = 2 * bj + (c & 1);
            const int gi = gi0 + ri, gj = gj0 + rj;
            const bool in = b < NB && gi >= 0 && gi < N && gj >= 0 && gj < N;
            inm[e] |= (in ? 1u : 0u) << c;
            yv[e][c] = in ? ys[(size_t)gi * N + gj] : 0.0;
            xc[e][c] = a * yv[e][c];
            xq[e][c] = 0.0;
            if (b < NB) X[(ri + 1) * RP + rj + 1] = xc[e][c];
        }
    }
    __syncthreads();
    for (int k = 1; k <= H; ++k) {
        const double om = A.om[k - 1], om1 = om - 1.0;
        const double* Xr = X + ((k - 1) & 1) * BUF;
        double* Xw = X + (k & 1) * BUF;
        // region rows / columns that must be exact after step k
        const int lo = HALO - (H - k), hi = HALO + TILE + (H - k);
#pragma unroll
        for (int e = 0; e < BPT; ++e) {
            const int b = threadIdx.x + G1_THREADS * e;
            const int r0 = 2 * (b / RB), c0 = 2 * (b % RB);
            if (b >= NB || r0 + 1 < lo || r0 >= hi || c0 + 1 < lo || c0 >= hi) continue;
            const int o = (r0 + 1) * RP + c0 + 1;  // LDS index of cell corner (r0, c0)
            const double up0 = Xr[o - RP], up1 = Xr[o - RP + 1];
            const double dn0 = Xr[o + 2 * RP], dn1 = Xr[o + 2 * RP + 1];
            const double lf0 = Xr[o - 1], lf1 = Xr[o + RP - 1];
            const double rt0 = Xr[o + 2], rt1 = Xr[o + RP + 2];
            const double x00 = xc[e][0], x01 = xc[e][1], x10 = xc[e][2], x11 = xc[e][3];
            double g[4];
            g[0] = 8.0 * x00 + (up1 + 2.0 * x01 + x11) + (lf1 + 2.0 * x10 + x11);
            g[1] = 8.0 * x01 + (up0 + 2.0 * x00 + x10) + (x10 + 2.0 * x11 + rt1);
            g[2] = 8.0 * x10 + (x01 + 2.0 * x11 + dn1) + (lf0 + 2.0 * x00 + x01);
            g[3] = 8.0 * x11 + (x00 + 2.0 * x10 + dn0) + (x00 + 2.0 * x01 + rt0);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double xo = xc[e][c];
                const double xn = ((inm[e] >> c) & 1) ? xo + om * (a * (yv[e][c] - g[c])) + om1 * (xo - xq[e][c]) : 0.0;
                xq[e][c] = xo;
                xc[e][c] = xn;
            }
            Xw[o] = xc[e][0];
            Xw[o + 1] = xc[e][1];
            Xw[o + RP] = xc[e][2];
            Xw[o + RP + 1] = xc[e][3];
        }
        __syncthreads();
    }
    // epilogue on the tile's interior, from the registers
#pragma unroll
    for (int e = 0; e < BPT; ++e) {
        const int b = threadIdx.x + G1_THREADS * e;
        if (b >= NB) continue;
        const int bi = b / RB, bj = b % RB;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int ri = 2 * bi + (c >> 1), rj = 2 * bj + (c & 1);
            const int gi = gi0 + ri, gj = gj0 + rj;
            if (ri < HALO || rj < HALO || ri >= HALO + TILE || rj >= HALO + TILE || gi >= N || gj >= N) continue;
            const size_t o = (size_t)s * N * N + (size_t)gi * N + gj;
            const double pv = xc[e][c];
            if (A.mode == 0) {
                A.out[o] = pv;
            } else if (A.mode == 1) {
                A.out[o] = A.a_out * pv;
            } else {
                const double xo = A.x[o];
                const double dx = A.xp ? (xo - A.xp[o]) : (xo - 0.0);
                A.out[o] = xo + A.w1 * (A.a_out * pv) + A.w0 * dx;
            }
        }
    }
}

template <int TILE, int HALO>
int launch_g1_cfg(const G1Args& A, int S, hipStream_t st) {
    const int N = 1 << A.n;
    const int T = (N + TILE - 1) / TILE;
    constexpr size_t lds = G1Cfg<TILE, HALO>::LDS;
    hipLaunchKernelGGL((g1_fused_kernel<TILE, HALO>), dim3(T, T, S), dim3(G1_THREADS), lds, st, A);
    return launch_check("g1_fused_kernel");
}

}  // namespace

int launch_g1_fused(int n, int deg, int mode, const double* y, const double* x, const double* xp, double a_out,
                    double w1, double w0, double* out, int S, hipStream_t st) {
    ADRT_REQUIRE(deg >= 1 && deg - 1 <= G1_MAXH, "g1_fused: g1_deg must be 1..17");
    ADRT_REQUIRE(n >= 1 && n <= 13, "g1_fused: need 1 <= n <= 13");
    if (S <= 0) return OK;
    G1Args A{};
    A.n = n;
    A.deg = deg;
    A.mode = mode;
    A.y = y;
    A.x = x;
    A.xp = xp;
    A.a_out = a_out;
    A.w1 = w1;
    A.w0 = w0;
    A.out = out;
    const double c = 12.0 / 20.0;
    double om = 1.0;
    for (int k = 1; k < deg; ++k) {
        om = (k == 1) ? 1.0 / (1.0 - c * c / 2.0) : 1.0 / (1.0 - c * c * om / 4.0);
        A.om[k - 1] = om;
    }
    const int N = 1 << n;
    // TILE 64 halves the halo overhead; TILE 32 gives a small batch more blocks
    const bool big = (long long)S * ((N + 63) / 64) * ((N + 63) / 64) >= 256;
    if (deg - 1 <= 14) return big ? launch_g1_cfg<64, 14>(A, S, st) : launch_g1_cfg<32, 14>(A, S, st);
    return big ? launch_g1_cfg<64, 16>(A, S, st) : launch_g1_cfg<32, 16>(A, S, st);
}

int launch_scale(const double* x, double alpha, double* out, long long count, hipStream_t st) {
    if (count <= 0) return OK;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(count)), dim3(IMG_THREADS), 0, st, x, alpha, out, count);
    return launch_check("scale_kernel");
}

int launch_cheb_update(const double* x, const double* xp, const double* r, double a, double w1, double w0,
                       double* out, long long count, hipStream_t st) {
    if (count <= 0) return OK;
    hipLaunchKernelGGL(cheb_update_kernel, dim3(grid_for(count)), dim3(IMG_THREADS), 0, st, x, xp, r, a, w1, w0, out,
                       count);
    return launch_check("cheb_update_kernel");
}

int launch_g1_step(int n, const double* y, const double* x, const double* xp, double a, double w1, double w0,
                   double* out, int S, hipStream_t st) {
    const long long count = (long long)S << (2 * n);
    if (count <= 0) return OK;
    ADRT_REQUIRE(n >= 1, "g1_step: need n >= 1");
    hipLaunchKernelGGL(g1_step_kernel, dim3(grid_for(count)), dim3(IMG_THREADS), 0, st, n, y, x, xp, a, w1, w0, out,
                       count);
    return launch_check("g1_step_kernel");
}

int launch_lorr(int n, const double* y, double* out, int S, hipStream_t st) {
    const long long count = (long long)S << (2 * n);
    if (count <= 0) return OK;
    ADRT_REQUIRE(n >= 1, "lorr: need n >= 1");
    hipLaunchKernelGGL(lorr_kernel, dim3(grid_for(count)), dim3(IMG_THREADS), 0, st, n, y, out, count);
    return launch_check("lorr_kernel");
}

}  // namespace adrt
