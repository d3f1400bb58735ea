// Image-space kernels of the outer iteration (fastops.py:247-258) and the level-1 preconditioners
// (fasthybrid.py:81-130): Chebyshev updates, the G_1 stencil step and the fused g1_inv, the lorr projector.  Images
// are double; the fused g1_inv iterates in float inside the FP32 outer steps of an FP32 plan.
#include <type_traits>

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
// (columns j-1..j+1) (fasthybrid.py:81-98); inside a 2 x 2 cell both partners are in the cell, so a step needs only
// the four neighbouring cells: the bottom row of the cell above, the top row of the cell below, the right column of
// the cell to the left and the left column of the cell to the right.  LDS holds the iterate as four planes, one per
// cell position (x00, x01, x10, x11), each indexed by cell: the lanes of a wave take consecutive cells, so every
// neighbour component they read is a run of consecutive words, free of bank conflicts (a cell-major 4-vector layout
// made each such read 4- to 8-way conflicted).  Double-buffered: step k reads buffer (k-1)&1 and writes k&1, one
// barrier per step.  Cells are numbered ring by ring: the tile's cells first,
// then the rings of halo cells at Chebyshev distance 1, 2, ... cells, and cell b belongs to thread b % THREADS.  The
// cells a step must compute (rings up to (deg - k) / 2) are then a prefix of that numbering, so whole waves drop out
// as the square shrinks.  Cells outside the image and a one-cell border around the region hold zero, the stencil's
// zero extension.  The epilogue optionally scales (mode 1) or applies the outer Chebyshev update (mode 2), so P and
// the outer step cost one pass.
constexpr int G1_MAXH = 16;
constexpr int G1_THREADS = 512;

struct G1Args {
    int n, deg, mode;
    int S;             // images
    const double* y;   // right-hand side (S, N, N)
    const float* y32;  // the same in float (float iteration only; y is then null)
    const double* x;   // outer iterate (mode 2)
    const double* xp;  // outer previous iterate (mode 2, may be null = 0)
    double a_out, w1, w0;
    double* out;
    double om[32];     // inner omega_k, k = 1..deg-1
    int lanes;         // mode 2 with per-image weights (lw1, lw0, lxp) instead of w1, w0 and xp for every image
    unsigned long long lxp;
    double lw1[G1_MAX_LANES], lw0[G1_MAX_LANES];
};


template <typename T, int TILE, int HALO>
struct G1Cfg {
    static constexpr int R = TILE + 2 * HALO;  // region side (even)
    static constexpr int RB = R / 2;           // 2 x 2 cells per region side
    static constexpr int TB = TILE / 2;        // cells per tile side
    static constexpr int HB = HALO / 2;        // halo rings (cells)
    // LDS pitch in cells: the region plus a zero border cell on each side, made odd so that a column of cells (the
    // halo rings' left and right sides, read two lanes per row) spreads over all banks instead of a few
    static constexpr int CP = (RB + 2) | 1;
    static constexpr int BUF = CP * CP;
    static constexpr int NB = RB * RB;
    static constexpr int BPT = (NB + G1_THREADS - 1) / G1_THREADS;
    static constexpr size_t LDS = 2 * 4 * (size_t)BUF * sizeof(T);  // two buffers of four planes
    // number of cells in the tile and its first j rings
    __host__ __device__ static constexpr int upto(int j) { return (TB + 2 * j) * (TB + 2 * j); }
};

// Region cell (row, column) of ring-ordered cell b and its ring (0 = inside the tile).
template <int TB, int HB>
__device__ __forceinline__ void ring_cell(int b, int& bi, int& bj, int& ring) {
    if (b < TB * TB) {
        bi = HB + b / TB;
        bj = HB + b % TB;
        ring = 0;
        return;
    }
    int j = 1, r = b - TB * TB;
    while (j < HB && r >= 4 * (TB + 2 * j - 1)) {
        r -= 4 * (TB + 2 * j - 1);
        ++j;
    }
    const int side = TB + 2 * j, o = HB - j;  // the ring is the border of a side x side square at (o, o)
    if (r < side) {
        bi = o;
        bj = o + r;
    } else if (r < 2 * side) {
        bi = o + side - 1;
        bj = o + r - side;
    } else {
        const int q = r - 2 * side;  // 2 (side - 2) cells of the left and right columns
        bi = o + 1 + (q >> 1);
        bj = (q & 1) ? o + side - 1 : o;
    }
    ring = j;
}

// T is the type of the inner iteration (double, or float inside the FP32 outer steps of an FP32 plan); the
// right-hand side, the outer iterates and the epilogue are double.
template <typename T, int TILE, int HALO>
__global__ __launch_bounds__(G1_THREADS) void g1_fused_kernel(G1Args A) {
    using C = G1Cfg<T, TILE, HALO>;
    constexpr int CP = C::CP, BUF = C::BUF, NB = C::NB, BPT = C::BPT, TB = C::TB, HB = C::HB;
    extern __shared__ __align__(16) unsigned char g1_smem[];
    T* X = reinterpret_cast<T*>(g1_smem);  // two buffers of four (CP, CP) planes: plane q of buffer u at X + (4u + q) BUF
    const int N = 1 << A.n;
    const int H = A.deg - 1;  // <= HALO
    // the tiles of one image on one XCD: neighbouring tiles' halos overlap, and are then read from that L2
    const int nt = (N + TILE - 1) / TILE;
    const int lb = xcd_block(blockIdx.x, nt * nt * A.S);
    if (lb < 0) return;
    const int s = lb / (nt * nt);
    const int gi0 = ((lb / nt) % nt) * TILE - HALO, gj0 = (lb % nt) * TILE - HALO;
    const size_t img = (size_t)s * N * N;
    // the epilogue's weights and previous iterate: the launch's, or this image's own (images at different steps)
    const bool hx = A.xp && (!A.lanes || ((A.lxp >> s) & 1ull));
    const double ew1 = A.lanes ? A.lw1[s] : A.w1, ew0 = A.lanes ? A.lw0[s] : A.w0;
    const T a = (T)(2.0 / 20.0);
    for (int i = threadIdx.x; i < 8 * BUF; i += G1_THREADS) X[i] = (T)0;
    __syncthreads();
    // float carries the increment d_k = x_k - x_(k-1) in xq and a y in yv: x_(k+1) = x_k + d_(k+1) with
    // d_(k+1) = w a (y - G_1 x_k) + (w - 1) d_k (the same recurrence, fewer instructions); double keeps the reference's
    // x + w (a r) + (w - 1)(x - x_prev) (fasthybrid.py:114-126) operation for operation.
    constexpr bool INCR = std::is_same<T, float>::value;
    T yv[BPT][4], xc[BPT][4], xq[BPT][4];
    unsigned inm[BPT];
    int lds[BPT];  // LDS cell index of the thread's cells
#pragma unroll
    for (int e = 0; e < BPT; ++e) {
        const int b = threadIdx.x + G1_THREADS * e;
        int bi = 0, bj = 0, ring = 0;
        if (b < NB) ring_cell<TB, HB>(b, bi, bj, ring);
        lds[e] = (bi + 1) * CP + bj + 1;
        inm[e] = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gi = gi0 + 2 * bi + (c >> 1), gj = gj0 + 2 * bj + (c & 1);
            const bool in = b < NB && gi >= 0 && gi < N && gj >= 0 && gj < N;
            inm[e] |= (in ? 1u : 0u) << c;
            const size_t gy = img + (size_t)gi * N + gj;
            yv[e][c] = in ? (A.y32 ? (T)A.y32[gy] : (T)A.y[gy]) : (T)0;
            xc[e][c] = a * yv[e][c];
            xq[e][c] = (T)0;
            if constexpr (INCR) {
                yv[e][c] = xc[e][c];  // a y
                xq[e][c] = xc[e][c];  // the first increment x_0 - 0
            }
        }
        if (b < NB) {
#pragma unroll
            for (int c = 0; c < 4; ++c) X[c * BUF + lds[e]] = xc[e][c];
        }
    }
    __syncthreads();
    // the outer update's x and x_prev of the tile's cells (the first TB^2 cells, slots e < EPI), loaded now so that
    // their latency hides behind the inner iteration
    constexpr int EPI = (TB * TB + G1_THREADS - 1) / G1_THREADS;
    double xo_[EPI][4], xpo_[EPI][4];
    if (A.mode == 2) {
#pragma unroll
        for (int e = 0; e < EPI; ++e) {
            const int b = threadIdx.x + G1_THREADS * e;
            const int bi = HB + b / TB, bj = HB + b % TB;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int gi = gi0 + 2 * bi + (c >> 1), gj = gj0 + 2 * bj + (c & 1);
                const bool in = b < TB * TB && gi < N && gj < N;
                const size_t o = (size_t)s * N * N + (size_t)gi * N + gj;
                xo_[e][c] = in ? A.x[o] : 0.0;
                xpo_[e][c] = (in && hx) ? A.xp[o] : 0.0;
            }
        }
    }
    for (int k = 1; k <= H; ++k) {
        const T om = (T)A.om[k - 1], om1 = (T)(A.om[k - 1] - 1.0);
        const T* Xr = X + ((k - 1) & 1) * 4 * BUF;
        T* Xw = X + (k & 1) * 4 * BUF;
        // the cells that must be exact after step k: rings j with 2 j - 1 <= H - k, a prefix of the numbering
        const int jmax = min((H - k + 1) / 2, HB);
        const int active = C::upto(jmax);
#pragma unroll
        for (int e = 0; e < BPT; ++e) {
            const int b = threadIdx.x + G1_THREADS * e;
            // a cell outside the image keeps x = 0 (its LDS entries stay as cleared); N and the region origin are
            // even, so a 2 x 2 cell is inside or outside as a whole
            if (b >= active || !inm[e]) continue;
            const int o = lds[e];
            const T up0 = Xr[2 * BUF + o - CP], up1 = Xr[3 * BUF + o - CP];  // x[r0 - 1][c0], x[r0 - 1][c0 + 1]
            const T dn0 = Xr[o + CP], dn1 = Xr[BUF + o + CP];                // x[r0 + 2][c0], x[r0 + 2][c0 + 1]
            const T lf0 = Xr[BUF + o - 1], lf1 = Xr[3 * BUF + o - 1];       // x[r0][c0 - 1], x[r0 + 1][c0 - 1]
            const T rt0 = Xr[o + 1], rt1 = Xr[2 * BUF + o + 1];              // x[r0][c0 + 2], x[r0 + 1][c0 + 2]
            const T x00 = xc[e][0], x01 = xc[e][1], x10 = xc[e][2], x11 = xc[e][3];
            // float: g = a (y - G_1 x); double: g = G_1 x
            T g[4];
            if constexpr (INCR) {
                // G_1 x at pixel p of the cell = 6 x_p + 2 S + (p's two neighbours outside the cell), S the cell's sum:
                // the double stencil below regrouped, 4 operations per pixel plus 3 for S instead of 6 per pixel
                const T S = (x00 + x01) + (x10 + x11);
                const T ma = -a, ma2 = (T)-2 * a, ma6 = (T)-6 * a;
                const T nb[4] = {up1 + lf1, up0 + rt1, dn1 + lf0, dn0 + rt0};
                const T xs[4] = {x00, x01, x10, x11};
#pragma unroll
                for (int c = 0; c < 4; ++c) g[c] = fma(ma6, xs[c], fma(ma2, S, fma(ma, nb[c], yv[e][c])));
            } else {
                const T two = (T)2, eight = (T)8;
                g[0] = eight * x00 + (up1 + two * x01 + x11) + (lf1 + two * x10 + x11);
                g[1] = eight * x01 + (up0 + two * x00 + x10) + (x10 + two * x11 + rt1);
                g[2] = eight * x10 + (x01 + two * x11 + dn1) + (lf0 + two * x00 + x01);
                g[3] = eight * x11 + (x00 + two * x10 + dn0) + (x00 + two * x01 + rt0);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if constexpr (INCR) {
                    const T d = om * g[c] + om1 * xq[e][c];
                    xq[e][c] = d;
                    xc[e][c] = xc[e][c] + d;
                } else {
                    const T xo = xc[e][c];
                    const T xn = xo + om * (a * (yv[e][c] - g[c])) + om1 * (xo - xq[e][c]);
                    xq[e][c] = xo;
                    xc[e][c] = xn;
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) Xw[c * BUF + o] = xc[e][c];
        }
        __syncthreads();
    }
    // epilogue on the tile's interior, from the registers
#pragma unroll
    for (int e = 0; e < BPT; ++e) {
        const int b = threadIdx.x + G1_THREADS * e;
        if (b >= TB * TB) continue;  // the tile's cells come first
        const int bi = HB + b / TB, bj = HB + b % TB;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int gi = gi0 + 2 * bi + (c >> 1), gj = gj0 + 2 * bj + (c & 1);
            if (gi >= N || gj >= N) continue;
            const size_t o = (size_t)s * N * N + (size_t)gi * N + gj;
            const double pv = (double)xc[e][c];
            if (A.mode == 0) {
                A.out[o] = pv;
            } else if (A.mode == 1) {
                A.out[o] = A.a_out * pv;
            } else if (e < EPI) {
                const double xo = xo_[e][c];
                const double dx = hx ? (xo - xpo_[e][c]) : (xo - 0.0);
                A.out[o] = xo + ew1 * (A.a_out * pv) + ew0 * dx;
            }
        }
    }
}

template <typename T, int TILE, int HALO>
int launch_g1_cfg(const G1Args& A, int S, hipStream_t st) {
    const int N = 1 << A.n;
    const int nt = (N + TILE - 1) / TILE;
    constexpr size_t lds = G1Cfg<T, TILE, HALO>::LDS;
    ADRT_REQUIRE((long long)nt * nt * S <= 0x7fffffffLL, "g1_fused: too many blocks");
    hipLaunchKernelGGL((g1_fused_kernel<T, TILE, HALO>), dim3(xcd_grid((long long)nt * nt * S)), dim3(G1_THREADS), lds, st,
                       A);
    return launch_check("g1_fused_kernel");
}

template <typename T>
int launch_g1_t(const G1Args& A, int S, int s_tile, hipStream_t st) {
    const int N = 1 << A.n;
    // TILE 64 halves the halo overhead; TILE 32 gives a small batch more blocks.  The float iteration always takes
    // TILE 64; the double one decides by s_tile, which a plan fixes (its max_batch).  A pixel's arithmetic then does
    // not depend on how many images share the launch, so a sinogram gives the same bits alone and inside a batch (the
    // two tile sizes differ by at most a rounding of the last bit).
    const bool big = std::is_same<T, float>::value || (long long)s_tile * ((N + 63) / 64) * ((N + 63) / 64) >= 256;
    if (A.deg - 1 <= 14) return big ? launch_g1_cfg<T, 64, 14>(A, S, st) : launch_g1_cfg<T, 32, 14>(A, S, st);
    return big ? launch_g1_cfg<T, 64, 16>(A, S, st) : launch_g1_cfg<T, 32, 16>(A, S, st);
}

}  // namespace

int launch_g1_fused(int n, int deg, int mode, const double* y, const double* x, const double* xp, double a_out,
                    double w1, double w0, double* out, int S, hipStream_t st, int prec, const float* y32, int s_tile,
                    const G1Lanes* lanes) {
    ADRT_REQUIRE(deg >= 1 && deg - 1 <= G1_MAXH, "g1_fused: g1_deg must be 1..17");
    ADRT_REQUIRE(!lanes || (mode == 2 && S <= G1_MAX_LANES), "g1_fused: per-image weights need mode 2 and S <= 64");
    ADRT_REQUIRE((y != nullptr) != (y32 != nullptr) && (y32 == nullptr || prec == PREC_F32),
                 "g1_fused: give y, or y32 for the float iteration");
    ADRT_REQUIRE(n >= 1 && n <= 13, "g1_fused: need 1 <= n <= 13");
    if (S <= 0) return OK;
    G1Args A{};
    A.n = n;
    A.deg = deg;
    A.mode = mode;
    A.y = y;
    A.y32 = y32;
    A.x = x;
    A.xp = xp;
    A.a_out = a_out;
    A.w1 = w1;
    A.w0 = w0;
    A.out = out;
    A.S = S;
    if (lanes) {
        A.lanes = 1;
        A.lxp = lanes->has_xp;
        for (int i = 0; i < S; ++i) {
            A.lw1[i] = lanes->w1[i];
            A.lw0[i] = lanes->w0[i];
        }
    }
    const double c = 12.0 / 20.0;
    double om = 1.0;
    for (int k = 1; k < deg; ++k) {
        om = (k == 1) ? 1.0 / (1.0 - c * c / 2.0) : 1.0 / (1.0 - c * c * om / 4.0);
        A.om[k - 1] = om;
    }
    const int stl = s_tile > 0 ? s_tile : S;
    return prec == PREC_F32 ? launch_g1_t<float>(A, S, stl, st) : launch_g1_t<double>(A, S, stl, st);
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
