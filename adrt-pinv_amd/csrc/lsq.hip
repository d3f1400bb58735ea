// Least-squares cross-check on the GPU: CGLS on min ||R x - b||_2 (pinv.py:23-48), batched over S sinograms.
//
// Every iteration is one forward ADRT and one transpose (the level kernels of adrt.hip, float64) plus vector updates.
// The per-sinogram scalars (gamma, alpha, beta, the stop flag) live on the device and are computed there from
// deterministic two-stage reductions (fixed block partition, fixed summation order), so a sinogram's result does not
// depend on the other sinograms of the batch.  A sinogram that meets its stop test ||R^T r_j|| <= tol ||R^T r_0||
// keeps alpha = 0 from then on, which freezes its x exactly where the reference's loop breaks.  The host reads the
// S stop flags once per iteration to end the loop when every sinogram has stopped.
#include <cmath>
#include <vector>

#include "../../include/adrt_hip.h"
#include "passes.hpp"

namespace adrt {

namespace {

constexpr int LSQ_THREADS = 256;
constexpr int LSQ_NB = 64;  // reduction blocks per sinogram
enum { SC_GAMMA = 0, SC_G0 = 1, SC_ALPHA = 2, SC_BETA = 3, SC_ACTIVE = 4, SC_ITERS = 5, SC_GNEW = 6, SC_N = 8 };

// partial[s][blk] = sum of v[s][i]^2 over the indices i = blk * 256 + t (mod LSQ_NB * 256), tree-reduced in LDS
__global__ __launch_bounds__(LSQ_THREADS) void dot_partial_kernel(const double* v, long long len, double* partial) {
    __shared__ double red[LSQ_THREADS];
    const int s = blockIdx.y;
    const double* vs = v + (size_t)s * len;
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * LSQ_THREADS + threadIdx.x; i < len; i += (long long)LSQ_NB * LSQ_THREADS)
        acc += vs[i] * vs[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = LSQ_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)s * LSQ_NB + blockIdx.x] = red[0];
}

// mode 0: gamma_0 = <s, s>; mode 1: alpha = gamma / <q, q>; mode 2: gnew = <s, s>, history, stop test, beta.
__global__ void scalars_kernel(int S, int mode, double tol, const double* partial, double* sc, double* hist,
                               int hist_len) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double sum = 0.0;
    for (int b = 0; b < LSQ_NB; ++b) sum += partial[(size_t)s * LSQ_NB + b];
    double* c = sc + (size_t)s * SC_N;
    if (mode == 0) {
        c[SC_GAMMA] = sum;
        c[SC_G0] = sqrt(sum);
        c[SC_ALPHA] = 0.0;
        c[SC_BETA] = 0.0;
        c[SC_ACTIVE] = sum > 0.0 ? 1.0 : 0.0;  // R^T b = 0: x = 0 is the least-squares solution
        c[SC_ITERS] = 0.0;
        if (hist_len > 0) hist[(size_t)s * hist_len] = sqrt(sum);
    } else if (mode == 1) {
        c[SC_ALPHA] = (c[SC_ACTIVE] != 0.0 && sum > 0.0) ? c[SC_GAMMA] / sum : 0.0;
    } else if (c[SC_ACTIVE] != 0.0) {
        const double it = c[SC_ITERS] + 1.0;
        c[SC_ITERS] = it;
        const double g = sqrt(sum);
        c[SC_GNEW] = g;
        if (hist_len > 0 && it < hist_len) hist[(size_t)s * hist_len + (long long)it] = g;
        if (g <= tol * c[SC_G0]) {
            c[SC_ACTIVE] = 0.0;
            c[SC_ALPHA] = 0.0;
        } else {
            c[SC_BETA] = sum / c[SC_GAMMA];
            c[SC_GAMMA] = sum;
        }
    } else {
        c[SC_ALPHA] = 0.0;
    }
}

// x += alpha p, r -= alpha q (per sinogram)
__global__ __launch_bounds__(LSQ_THREADS) void update_xr_kernel(double* x, const double* p, double* r, const double* q,
                                                                const double* sc, long long nx, long long nr) {
    const int s = blockIdx.y;
    const double a = sc[(size_t)s * SC_N + SC_ALPHA];
    if (a == 0.0) return;
    const long long n = nx > nr ? nx : nr;
    for (long long i = (long long)blockIdx.x * LSQ_THREADS + threadIdx.x; i < n; i += (long long)gridDim.x * LSQ_THREADS) {
        if (i < nx) x[(size_t)s * nx + i] += a * p[(size_t)s * nx + i];
        if (i < nr) r[(size_t)s * nr + i] -= a * q[(size_t)s * nr + i];
    }
}

// p = s + beta p for the sinograms still running
__global__ __launch_bounds__(LSQ_THREADS) void update_p_kernel(double* p, const double* sv, const double* sc,
                                                               long long nx) {
    const int s = blockIdx.y;
    if (sc[(size_t)s * SC_N + SC_ACTIVE] == 0.0) return;
    const double beta = sc[(size_t)s * SC_N + SC_BETA];
    for (long long i = (long long)blockIdx.x * LSQ_THREADS + threadIdx.x; i < nx; i += (long long)gridDim.x * LSQ_THREADS)
        p[(size_t)s * nx + i] = sv[(size_t)s * nx + i] + beta * p[(size_t)s * nx + i];
}

// r = b - q
__global__ __launch_bounds__(LSQ_THREADS) void sub_kernel(double* r, const double* b, const double* q, long long count) {
    for (long long i = (long long)blockIdx.x * LSQ_THREADS + threadIdx.x; i < count; i += (long long)gridDim.x * LSQ_THREADS)
        r[i] = b[i] - q[i];
}

unsigned blocks_for(long long count) {
    long long b = (count + LSQ_THREADS - 1) / LSQ_THREADS;
    return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

struct CglsWork {
    double *r, *q, *aw, *s, *p, *partial, *sc, *hist;
};

CglsWork carve(void* work, int n, int S, int hist_len) {
    const size_t N = (size_t)1 << n, Mb = 4 * (2 * N - 1) * N, NN = N * N;
    double* w = (double*)work;
    CglsWork c{};
    c.r = w;
    c.q = c.r + S * Mb;
    c.aw = c.q + S * Mb;  // ADRT scratch: two sinogram-sized arrays (the transpose ping-pongs)
    c.s = c.aw + 2 * S * Mb;
    c.p = c.s + S * NN;
    c.partial = c.p + S * NN;
    c.sc = c.partial + (size_t)S * LSQ_NB;
    c.hist = c.sc + (size_t)S * SC_N;
    return c;
}

}  // namespace

}  // namespace adrt

using namespace adrt;

extern "C" {

unsigned long long adrt_cgls_work_bytes(int n, int S, int hist_len) {
    if (n < 1 || n > 13 || S < 0 || hist_len < 0) return 0;
    const size_t N = (size_t)1 << n, Mb = 4 * (2 * N - 1) * N, NN = N * N;
    return sizeof(double) * ((size_t)S * (4 * Mb + 2 * NN + LSQ_NB + SC_N) + (size_t)S * hist_len);
}

int adrt_cgls(int n, const double* d_b, double* d_x, int S, double tol, int maxit, int x0, void* d_work, int* iters,
              double* hist, int hist_len, void* stream) {
    return adrt_cgls_cb(n, d_b, d_x, S, tol, maxit, x0, d_work, iters, hist, hist_len, stream, nullptr, nullptr);
}

int adrt_cgls_cb(int n, const double* d_b, double* d_x, int S, double tol, int maxit, int x0, void* d_work,
                 int* iters, double* hist, int hist_len, void* stream, adrt_cgls_callback callback, void* user) {
    ADRT_REQUIRE(n >= 1 && n <= 13, "cgls: need 1 <= n <= 13");
    ADRT_REQUIRE(S >= 0 && maxit >= 0 && hist_len >= 0, "cgls: negative S, maxit or hist_len");
    ADRT_REQUIRE(tol >= 0.0, "cgls: tol must be >= 0");
    if (S == 0) return OK;
    ADRT_REQUIRE(d_b && d_x && d_work && iters, "cgls: null buffer");
    ADRT_REQUIRE(hist_len == 0 || hist, "cgls: hist_len > 0 needs a history buffer");
    hipStream_t st = (hipStream_t)stream;
    const long long N = 1LL << n, Mb = 4 * (2 * N - 1) * N, NN = N * N;
    CglsWork w = carve(d_work, n, S, hist_len);
    const dim3 gx(blocks_for(NN), S), gr(blocks_for(NN > Mb ? NN : Mb), S), gd(LSQ_NB, S);
    const unsigned gs = (unsigned)((S + 63) / 64);
    auto dot = [&](const double* v, long long len, int mode) -> int {
        hipLaunchKernelGGL(dot_partial_kernel, gd, dim3(LSQ_THREADS), 0, st, v, len, w.partial);
        ADRT_TRY(launch_check("dot_partial_kernel"));
        hipLaunchKernelGGL(scalars_kernel, dim3(gs), dim3(64), 0, st, S, mode, tol, w.partial, w.sc, w.hist, hist_len);
        return launch_check("scalars_kernel");
    };
    // r = b - R x, s = R^T r, p = s, gamma = <s, s>  (pinv.py:27-33)
    if (x0) {
        ADRT_TRY(launch_adrt_forward(n, 2, d_x, w.q, w.aw, S, st));
        hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(S * Mb)), dim3(LSQ_THREADS), 0, st, w.r, d_b, w.q, S * Mb);
        ADRT_TRY(launch_check("sub_kernel"));
    } else {
        ADRT_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * S * NN, st));
        ADRT_HIP(hipMemcpyAsync(w.r, d_b, sizeof(double) * S * Mb, hipMemcpyDeviceToDevice, st));
    }
    ADRT_TRY(launch_adrt_transpose(n, 2, w.r, w.s, w.aw, S, st));
    ADRT_HIP(hipMemcpyAsync(w.p, w.s, sizeof(double) * S * NN, hipMemcpyDeviceToDevice, st));
    ADRT_TRY(dot(w.s, NN, 0));
    std::vector<double> sc((size_t)S * SC_N);
    auto any_active = [&](bool& any) -> int {
        ADRT_HIP(hipMemcpyAsync(sc.data(), w.sc, sizeof(double) * sc.size(), hipMemcpyDeviceToHost, st));
        ADRT_HIP(hipStreamSynchronize(st));
        any = false;
        for (int s = 0; s < S; ++s) any |= sc[(size_t)s * SC_N + SC_ACTIVE] != 0.0;
        return OK;
    };
    bool any = true;
    ADRT_TRY(any_active(any));
    std::vector<int> ran(callback ? S : 0);
    std::vector<double> gnorm(callback ? S : 0);
    for (int it = 0; it < maxit && any; ++it) {
        // the sinograms that run this iteration: those still active after the last one (pinv.py:35-47)
        for (size_t s = 0; s < ran.size(); ++s) ran[s] = sc[s * SC_N + SC_ACTIVE] != 0.0;
        ADRT_TRY(launch_adrt_forward(n, 2, w.p, w.q, w.aw, S, st));  // q = R p
        ADRT_TRY(dot(w.q, Mb, 1));                                     // alpha = gamma / <q, q>
        hipLaunchKernelGGL(update_xr_kernel, gr, dim3(LSQ_THREADS), 0, st, d_x, w.p, w.r, w.q, w.sc, NN, Mb);
        ADRT_TRY(launch_check("update_xr_kernel"));
        ADRT_TRY(launch_adrt_transpose(n, 2, w.r, w.s, w.aw, S, st));  // s = R^T r
        ADRT_TRY(dot(w.s, NN, 2));                                      // gnew, stop test, beta
        hipLaunchKernelGGL(update_p_kernel, gx, dim3(LSQ_THREADS), 0, st, w.p, w.s, w.sc, NN);
        ADRT_TRY(launch_check("update_p_kernel"));
        ADRT_TRY(any_active(any));
        if (callback) {
            // after the stop test, before the next iteration: x and ||R^T r|| of iteration it (pinv.py:43-44)
            for (int s = 0; s < S; ++s) gnorm[s] = ran[s] ? sc[(size_t)s * SC_N + SC_GNEW] : 0.0;
            callback(it, d_x, gnorm.data(), ran.data(), S, user);
        }
    }
    ADRT_TRY(any_active(any));
    for (int s = 0; s < S; ++s) iters[s] = (int)sc[(size_t)s * SC_N + SC_ITERS];
    if (hist_len > 0) {
        ADRT_HIP(hipMemcpyAsync(hist, w.hist, sizeof(double) * S * hist_len, hipMemcpyDeviceToHost, st));
        ADRT_HIP(hipStreamSynchronize(st));
    }
    return OK;
}

}  // extern "C"
