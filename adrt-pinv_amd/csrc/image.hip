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

// g1_inv fused into one launch: each block owns a TILE x TILE output tile and iterates the Chebyshev solve of G_1 on
// the tile grown by a halo of (deg - 1) pixels, so every output pixel sees exact neighbours through all steps (the
// stencil reaches one pixel per step).  The current iterate lives in LDS (neighbour reads); y, the previous iterate
// and the new values stay in registers.  Pixels outside the image are held at zero like the stencil's zero extension.
// The epilogue optionally scales (mode 1) or applies the outer Chebyshev update (mode 2) so P and the outer step cost
// one pass over the image.
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

// Region slot e of a thread is (ri, rj) = divmod(tid + 512 e, REG); the region of a launch is the top-left R x R of
// the REG x REG slot square (R = TILE + 2 H, REG = TILE + 32).  TILE 64 halves the halo overhead; TILE 32 gives a
// small batch more blocks.
template <int G1_TILE>
__global__ __launch_bounds__(G1_THREADS) void g1_fused_kernel(G1Args A) {
    constexpr int G1_REG = G1_TILE + 2 * G1_MAXH;
    constexpr int G1_RP = G1_REG + 2;  // LDS row pitch: region plus a zero border
    constexpr int G1_PER = G1_REG * G1_REG / G1_THREADS;
    static_assert(G1_REG * G1_REG % G1_THREADS == 0, "region slots must tile the block");
    extern __shared__ double X[];  // G1_RP * G1_RP
    const int N = 1 << A.n;
    const int H = A.deg - 1;
    const int R = G1_TILE + 2 * H;
    const int s = blockIdx.z;
    const int ti0 = blockIdx.y * G1_TILE - H, tj0 = blockIdx.x * G1_TILE - H;
    const double* ys = A.y + (size_t)s * N * N;
    const double a = 2.0 / 20.0;
    for (int i = threadIdx.x; i < G1_RP * G1_RP; i += G1_THREADS) X[i] = 0.0;
    __syncthreads();
    double yv[G1_PER], xp[G1_PER];  // the current iterate lives in LDS only
    unsigned inmask = 0;
#pragma unroll
    for (int e = 0; e < G1_PER; ++e) {
        const int idx = threadIdx.x + G1_THREADS * e;
        const int ri = idx / G1_REG, rj = idx % G1_REG;
        const int gi = ti0 + ri, gj = tj0 + rj;
        const bool in = ri < R && rj < R && gi >= 0 && gi < N && gj >= 0 && gj < N;
        inmask |= (in ? 1u : 0u) << e;
        yv[e] = in ? ys[(size_t)gi * N + gj] : 0.0;
        xp[e] = 0.0;
        X[(ri + 1) * G1_RP + rj + 1] = a * yv[e];
    }
    __syncthreads();
    for (int k = 1; k < A.deg; ++k) {
        const double om = A.om[k - 1];
        double xn[G1_PER];
#pragma unroll
        for (int e = 0; e < G1_PER; ++e) {
            const int idx = threadIdx.x + G1_THREADS * e;
            const int ri = idx / G1_REG, rj = idx % G1_REG;
            // parity of the global index equals the parity of the region index shifted by the (even) origin offset
            const int pi = (ti0 + ri) & 1, pj = (tj0 + rj) & 1;
            // pixels outside the image hold 0 in X and the border row/column is 0, so every neighbour read is plain
            const int c = (ri + 1) * G1_RP + rj + 1;
            const int cj = c + (pj ? -1 : 1), ci = c + (pi ? -G1_RP : G1_RP);
            const double xk = X[c];
            const double t1 = 2.0 * X[cj] + X[cj - G1_RP] + X[cj + G1_RP];
            const double t2 = 2.0 * X[ci] + X[ci - 1] + X[ci + 1];
            const double g = 8.0 * xk + t1 + t2;
            const double r = yv[e] - g;
            xn[e] = ((inmask >> e) & 1) ? xk + om * (a * r) + (om - 1.0) * (xk - xp[e]) : 0.0;
            xp[e] = xk;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < G1_PER; ++e) {
            const int idx = threadIdx.x + G1_THREADS * e;
            const int ri = idx / G1_REG, rj = idx % G1_REG;
            X[(ri + 1) * G1_RP + rj + 1] = xn[e];
        }
        __syncthreads();
    }
    // epilogue on the tile's interior
#pragma unroll
    for (int e = 0; e < G1_PER; ++e) {
        const int idx = threadIdx.x + G1_THREADS * e;
        const int ri = idx / G1_REG, rj = idx % G1_REG;
        const int gi = ti0 + ri, gj = tj0 + rj;
        if (ri < H || rj < H || ri >= H + G1_TILE || rj >= H + G1_TILE || gi >= N || gj >= N) continue;
        const size_t o = (size_t)s * N * N + (size_t)gi * N + gj;
        const double pv = X[(ri + 1) * G1_RP + rj + 1];
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
    const long long tiles64 = (long long)S * ((N + 63) / 64) * ((N + 63) / 64);
    if (tiles64 >= 256) {
        const int T = (N + 63) / 64;
        constexpr int RP = 64 + 2 * G1_MAXH + 2;
        hipLaunchKernelGGL(g1_fused_kernel<64>, dim3(T, T, S), dim3(G1_THREADS), sizeof(double) * RP * RP, st, A);
    } else {
        const int T = (N + 31) / 32;
        constexpr int RP = 32 + 2 * G1_MAXH + 2;
        hipLaunchKernelGGL(g1_fused_kernel<32>, dim3(T, T, S), dim3(G1_THREADS), sizeof(double) * RP * RP, st, A);
    }
    return launch_check("g1_fused_kernel");
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
