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

}  // namespace

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
