// Dense low-frequency rows as a batched complex GEMM.
//
// For a frequency k whose degree exceeds dense_cap, the reference applies the exact Tikhonov filter F_k inside the op
// list (fastops.py:28-39, 104).  Every op of a packed pass is linear and row-local, so the whole op list of such a row
// is one N x N matrix W_k (e.g. B1^H M' F B1 for G_h), built once at plan creation.  A pass then multiplies the
// 2S rows of frequency k (both packed pairs of every sinogram) by W_k: out_r = W_k in_r.  One block owns TR rows of
// one frequency and all N outputs, so it may write its rows in place; W_k^T is streamed from L2/HBM (coalesced over
// the outputs) and reused by the block's TR rows, whose inputs are staged through LDS in chunks of TJ columns.
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
// C[r][i] = sum_j Y[r][j] WT[j][i] for the SP rows r of one dense frequency, complex, as four real
// v_mfma_f64_16x16x4 products per k-step (C_re += Yr Wr - Yi Wi, C_im += Yr Wi + Yi Wr).  A block computes a 64 x 64
// output tile with four waves of 32 x 32 (2 x 2 MFMA tiles); k runs in chunks of 16 staged in LDS.  Operand layouts of
// the 16x16x4 f64 instruction (checked on gfx950): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, and
// D[4 r + lane / 16][lane % 16] in accumulator element r.
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int MF_BM = 64, MF_BN = 64, MF_KC = 16;

template <typename Tg>
__global__ __launch_bounds__(256) void dense_mfma_kernel(int n, int L, int SP, const int* dense_k, const c128* WT,
                                                         const void* Yin, void* Zout) {
    __shared__ c128 As[MF_BM][MF_KC + 1];
    __shared__ c128 Bs[MF_KC][MF_BN + 1];
    const int N = 1 << n;
    const int slot = blockIdx.z;
    const int k = dense_k[slot];
    const int r0 = blockIdx.y * MF_BM, c0 = blockIdx.x * MF_BN;
    const c128* W = WT + (size_t)slot * N * N;
    const cplx<Tg>* Y = reinterpret_cast<const cplx<Tg>*>(Yin);
    cplx<Tg>* Z = reinterpret_cast<cplx<Tg>*>(Zout);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    v4d cre[2][2], cim[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) cre[a][b] = cim[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < N; j0 += MF_KC) {
        for (int idx = threadIdx.x; idx < MF_BM * MF_KC; idx += 256) {
            const int rr = idx / MF_KC, jj = idx % MF_KC;
            const int sp = r0 + rr;
            As[rr][jj] = sp < SP ? cast<double>(Y[((size_t)sp * L + k) * N + j0 + jj]) : mk(0.0, 0.0);
        }
        for (int idx = threadIdx.x; idx < MF_KC * MF_BN; idx += 256) {
            const int jj = idx / MF_BN, cc = idx % MF_BN;
            Bs[jj][cc] = (c0 + cc < N) ? W[(size_t)(j0 + jj) * N + c0 + cc] : mk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < MF_KC; kk += 4) {
            double are[2], aim[2], bre[2], bim[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const c128 v = As[wr + a * 16 + (lane & 15)][kk + (lane >> 4)];
                are[a] = v.x;
                aim[a] = v.y;
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const c128 v = Bs[kk + (lane >> 4)][wc + b * 16 + (lane & 15)];
                bre[b] = v.x;
                bim[b] = v.y;
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    cre[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[a], bre[b], cre[a][b], 0, 0, 0);
                    cre[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-aim[a], bim[b], cre[a][b], 0, 0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(are[a], bim[b], cim[a][b], 0, 0, 0);
                    cim[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(aim[a], bre[b], cim[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int sp = r0 + wr + a * 16 + 4 * e + (lane >> 4);
                const int i = c0 + wc + b * 16 + (lane & 15);
                if (sp < SP && i < N) Z[((size_t)sp * L + k) * N + i] = cast<Tg>(mk(cre[a][b][e], cim[a][b][e]));
            }
}

// copies the dense rows of Y2 back into Y
template <typename Tg>
__global__ void dense_copy_kernel(int n, int L, int SP, const int* dense_k, const void* src, void* dst) {
    const int N = 1 << n;
    const int slot = blockIdx.y;
    const int k = dense_k[slot];
    const cplx<Tg>* S = reinterpret_cast<const cplx<Tg>*>(src);
    cplx<Tg>* D = reinterpret_cast<cplx<Tg>*>(dst);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)SP * N;
         e += (long long)gridDim.x * blockDim.x) {
        const size_t o = ((size_t)(e / N) * L + k) * N + (e % N);
        D[o] = S[o];
    }
}

template <typename Tg>
int launch_mfma(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                void* tmp, hipStream_t st) {
    const int N = 1 << n;
    void* out = (Z == Y) ? tmp : Z;
    hipLaunchKernelGGL(dense_mfma_kernel<Tg>, dim3((N + MF_BN - 1) / MF_BN, (SP + MF_BM - 1) / MF_BM, n_dense),
                       dim3(256), 0, st, n, L, SP, dense_k, WT, Y, out);
    ADRT_TRY(launch_check("dense_mfma_kernel"));
    if (out != Z) {
        long long cnt = (long long)SP * N;
        unsigned gx = (unsigned)((cnt + 255) / 256 < 4096 ? (cnt + 255) / 256 : 4096);
        hipLaunchKernelGGL(dense_copy_kernel<Tg>, dim3(gx, n_dense), dim3(256), 0, st, n, L, SP, dense_k, out, Z);
        ADRT_TRY(launch_check("dense_copy_kernel"));
    }
    return OK;
}

}  // namespace

// Z_r = W_slot Y_r on the FP64 matrix cores (requires N >= 16 and, when Z aliases Y, a row buffer tmp of Y's size).
int launch_dense_mfma(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                      void* tmp, int prec, hipStream_t st) {
    if (n_dense <= 0 || SP <= 0) return OK;
    ADRT_REQUIRE(n >= 4 && n <= 11, "dense_mfma: need 4 <= n <= 11");
    ADRT_REQUIRE(Z != Y || tmp, "dense_mfma: in-place use needs a scratch row buffer");
    if (prec == PREC_F64) return launch_mfma<double>(n, L, SP, n_dense, dense_k, WT, Y, Z, tmp, st);
    return launch_mfma<float>(n, L, SP, n_dense, dense_k, WT, Y, Z, tmp, st);
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
