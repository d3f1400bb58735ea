// Dense low-frequency rows as a batched complex GEMM.
//
// For a frequency k whose degree exceeds dense_cap, the reference applies the exact Tikhonov filter F_k inside the op
// list (fastops.py:28-39, 104).  Every op of a packed pass is linear and row-local, so the whole op list of such a row
// is one N x N matrix W_k (e.g. B1^H M' F B1 for G_h), built once at plan creation.  A pass then multiplies the
// 2S rows of frequency k (both packed pairs of every sinogram) by W_k: out_r = W_k in_r.  One block owns TR rows of
// one frequency and all N outputs, so it may write its rows in place; W_k^T is streamed from L2/HBM (coalesced over
// the outputs) and reused by the block's TR rows, whose inputs are staged through LDS in chunks of TJ columns.
#include "dense.hpp"
#include "rows.hpp"

namespace adrt {

namespace {

constexpr int DG_THREADS = 256;
constexpr int DG_TJ = 32;

template <typename Tg, int PER, int TR>
__global__ __launch_bounds__(DG_THREADS) void dense_gemm_kernel(int n, int L, int SP, const int* dense_k,
                                                                const c128* WT, const void* Yin, void* Yout) {
    __shared__ c128 ys[TR][DG_TJ];
    const int N = 1 << n;
    const int slot = blockIdx.y;
    const int k = dense_k[slot];
    const int sp0 = blockIdx.x * TR;
    const c128* W = WT + (size_t)slot * N * N;
    const cplx<Tg>* Y = reinterpret_cast<const cplx<Tg>*>(Yin);
    cplx<Tg>* Z = reinterpret_cast<cplx<Tg>*>(Yout);
    c128 acc[TR][PER];
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
        for (int e = 0; e < PER; ++e) acc[r][e] = mk(0.0, 0.0);
    for (int j0 = 0; j0 < N; j0 += DG_TJ) {
        for (int idx = threadIdx.x; idx < TR * DG_TJ; idx += DG_THREADS) {
            const int r = idx / DG_TJ, jj = idx % DG_TJ;
            const int sp = sp0 + r;
            ys[r][jj] = (sp < SP && j0 + jj < N) ? cast<double>(Y[((size_t)sp * L + k) * N + j0 + jj]) : mk(0.0, 0.0);
        }
        __syncthreads();
        const int jn = (N - j0) < DG_TJ ? (N - j0) : DG_TJ;
        // eight columns of W^T in flight per step (the loop is otherwise bound by load latency)
        for (int jb = 0; jb < jn; jb += 8) {
            c128 w[8][PER];
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int e = 0; e < PER; ++e) {
                    const int i = threadIdx.x + DG_THREADS * e;
                    w[q][e] = (i < N && jb + q < jn) ? W[(size_t)(j0 + jb + q) * N + i] : mk(0.0, 0.0);
                }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
#pragma unroll
                for (int r = 0; r < TR; ++r) {
                    const c128 yv = ys[r][jb + q < DG_TJ ? jb + q : 0];
#pragma unroll
                    for (int e = 0; e < PER; ++e) acc[r][e] = add(acc[r][e], mul(w[q][e], yv));
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < TR; ++r) {
        const int sp = sp0 + r;
        if (sp >= SP) continue;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int i = threadIdx.x + DG_THREADS * e;
            if (i < N) Z[((size_t)sp * L + k) * N + i] = cast<Tg>(acc[r][e]);
        }
    }
}

template <typename Tg, int PER, int TR>
int launch_t(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
             hipStream_t st) {
    const unsigned gx = (unsigned)((SP + TR - 1) / TR);
    hipLaunchKernelGGL((dense_gemm_kernel<Tg, PER, TR>), dim3(gx, n_dense), dim3(DG_THREADS), 0, st, n, L, SP, dense_k,
                       WT, Y, Z);
    return launch_check("dense_gemm_kernel");
}

template <typename Tg>
int launch_n(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
             hipStream_t st) {
    // PER outputs per thread x TR rows per block; few rows (one sinogram) get a narrow TR so no block idles
    const bool few = SP <= 4;
    if (n <= 8)
        return few ? launch_t<Tg, 1, 2>(n, L, SP, n_dense, dense_k, WT, Y, Z, st)
                   : launch_t<Tg, 1, 16>(n, L, SP, n_dense, dense_k, WT, Y, Z, st);
    if (n == 9)
        return few ? launch_t<Tg, 2, 2>(n, L, SP, n_dense, dense_k, WT, Y, Z, st)
                   : launch_t<Tg, 2, 8>(n, L, SP, n_dense, dense_k, WT, Y, Z, st);
    if (n == 10)
        return few ? launch_t<Tg, 4, 2>(n, L, SP, n_dense, dense_k, WT, Y, Z, st)
                   : launch_t<Tg, 4, 4>(n, L, SP, n_dense, dense_k, WT, Y, Z, st);
    return launch_t<Tg, 8, 2>(n, L, SP, n_dense, dense_k, WT, Y, Z, st);
}

// ---------------------------------------------------------------- MFMA version (FP64 matrix cores)
// Each wave owns a block of RT x CT 16 x 16 output tiles (dense.hpp: dense::tile).  A block is four waves: WR of them
// on consecutive row blocks and 4 / WR on consecutive output blocks (WR = 4 with many rows, fewer with few), so with
// many rows the block's W^T columns are fetched once into L1 and shared, and with few rows (a single sinogram) every
// wave still has rows to work on.  Blocks are numbered so that the blocks of one frequency run on one XCD (the
// dispatcher deals consecutive blocks to the 8 XCDs in turn): the frequency's W^T then comes from HBM once per XCD
// and from that XCD's L2 for its other row groups.
// tiles per wave of the stand-alone kernel (rows, outputs).  Measured at N = 256 / 512 with 128 / 92 rows per
// frequency (us per launch): 1 x 1 with 8 k-steps per load group 67 / 398; 1 x 2 62 / 343; 2 x 1 73 / 380; 2 x 2 79 / 352;
// 4 x 1 88 / 474.  The four waves of a block already share W through L1, so row blocking only costs registers.
constexpr int SA_RT = 1, SA_CT = 2;

template <typename Tg, int U, int RT, int CT>
__global__ __launch_bounds__(256, 2) void dense_mfma_kernel(int n, DenseArgs D, const void* Yin, void* Zout, int gx,
                                                            int gy, int nblk, int wr) {
    const int N = 1 << n;
    const int lb = xcd_block(blockIdx.x, nblk);
    if (lb < 0) return;  // (the kernel has no barrier)
    const int bx = lb % gx, by = (lb / gx) % gy, slot = lb / (gx * gy);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = (by * wr + wave % wr) * 16 * RT;
    const int c0 = (bx * (4 / wr) + wave / wr) * 16 * CT;
    if (r0 >= dense::unit_rows(D) || c0 >= N) return;
    dense::tile<Tg, U, RT, CT>(n, D, slot, r0, c0, lane, Yin, Zout);
}

template <typename Tg, int U, int RT, int CT>
int launch_mfma_rc(int n, const DenseArgs& D, const void* Y, void* Z, hipStream_t st) {
    const int N = 1 << n;
    const int rb = (dense::unit_rows(D) + 16 * RT - 1) / (16 * RT);  // row blocks
    const int cb = (N / 16 + CT - 1) / CT;              // output blocks
    const int wr = rb >= 4 ? 4 : (rb >= 2 ? 2 : 1);     // waves of a block along the rows
    const int gx = (cb + 4 / wr - 1) / (4 / wr), gy = (rb + wr - 1) / wr;
    const long long nblk = (long long)gx * gy * D.n_dense;
    ADRT_REQUIRE(nblk <= 0x7fffffffLL, "dense_mfma: too many blocks");
    hipLaunchKernelGGL((dense_mfma_kernel<Tg, U, RT, CT>), dim3(xcd_grid(nblk)), dim3(256), 0, st, n, D, Y, Z, gx, gy,
                       (int)nblk, wr);
    return launch_check("dense_mfma_kernel");
}

template <typename Tg>
int launch_mfma(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                const int* pair_k, const int* wslot, hipStream_t st) {
    DenseArgs D;
    D.dense_k = dense_k;
    D.pair_k = pair_k;
    D.wslot = wslot;
    D.WT = WT;
    D.n_dense = n_dense;
    D.L = L;
    D.SP = SP;
    D.rows = pair_k ? 2 * SP : SP;
    SP = D.rows;
    // blocked waves beyond one row tile; a single row tile (one or a few sinograms, where W streams from HBM) keeps
    // 16 x 16 tiles, twice the waves, with deeper load groups (8 k-steps for complex64 rows; complex128 rows take 4:
    // the larger group's registers would cost half the waves)
    if (SP > 16) return launch_mfma_rc<Tg, 4, SA_RT, SA_CT>(n, D, Y, Z, st);
    if ((1 << n) >= 32 && sizeof(Tg) == 4) return launch_mfma_rc<Tg, 8, 1, 1>(n, D, Y, Z, st);
    return launch_mfma_rc<Tg, 4, 1, 1>(n, D, Y, Z, st);
}
}  // namespace

// Z_r = W_slot Y_r on the FP64 matrix cores (requires N >= 16 and Z distinct from Y); with pair_k, units of
// conjugate frequency pairs (DenseArgs).
int launch_dense_mfma(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                      int prec, hipStream_t st, const int* pair_k, const int* wslot) {
    if (n_dense <= 0 || SP <= 0) return OK;
    ADRT_REQUIRE(n >= 4 && n <= 11, "dense_mfma: need 4 <= n <= 11");
    ADRT_REQUIRE(Z != Y, "dense_mfma: the output must not alias the input rows");
    if (prec == PREC_F64) return launch_mfma<double>(n, L, SP, n_dense, dense_k, WT, Y, Z, pair_k, wslot, st);
    return launch_mfma<float>(n, L, SP, n_dense, dense_k, WT, Y, Z, pair_k, wslot, st);
}

// Z_r = W_slot Y_r for the SP rows of every dense frequency (rows sp * L + k); Z may alias Y.
int launch_dense_gemm(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                      int prec, hipStream_t st) {
    if (n_dense <= 0 || SP <= 0) return OK;
    ADRT_REQUIRE(n >= 1 && n <= 11, "dense_gemm: need 1 <= n <= 11");
    if (prec == PREC_F64) return launch_n<double>(n, L, SP, n_dense, dense_k, WT, Y, Z, st);
    return launch_n<float>(n, L, SP, n_dense, dense_k, WT, Y, Z, st);
}

}  // namespace adrt
