// One 16 x 16 output tile of the dense low-frequency GEMM on the FP64 matrix cores, shared by the stand-alone dense
// kernel (dense.hip) and the dense workers of the register row kernel (rows_reg.hpp).
//
// C[r][i] = sum_j Y[r][j] WT[j][i] for 16 rows r of one dense frequency and 16 outputs i, complex, as three real
// v_mfma_f64_16x16x4 products per k-step (Gauss: P1 += Yr Wr, P2 += Yi Wi, P3 += (Yr + Yi)(Wr + Wi), then
// C_re = P1 - P2, C_im = P3 - P1 - P2; the sums Yr + Yi of float rows are exact in double).  Operand layouts of the
// 16x16x4 f64 instruction (checked on gfx950): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, and
// D[4 r + lane / 16][lane % 16] in accumulator element r.  The k-steps run in ascending order whatever U is, so
// every caller gets the same bits for a row.
#pragma once
#include "common.hpp"

namespace adrt {

// The dense rows of a pass: rows sp * L + dense_k[slot] (sp < SP) are multiplied by W_slot (W^T stored, N x N).
struct DenseArgs {
    const int* dense_k = nullptr;
    const c128* WT = nullptr;
    int n_dense = 0;
    int L = 0;
    int SP = 0;
    int workers = 0;  // register row kernel: waves that take dense tiles (0 = none)
    int dblk = 0;     // register row kernel: leading blocks that hold those waves (a multiple of 8)
};

namespace dense {

typedef double v4d __attribute__((ext_vector_type(4)));

// The tile of rows r0..r0+15 and outputs c0..c0+15 of dense slot `slot`, computed by one wave (lane 0..63).  U
// k-steps (of 4) per load group, the next group's loads in flight while the current group's 3 U MFMAs run.  Zout
// must not alias Yin (other waves still read the frequency's rows).
template <typename Tg, int U>
__device__ __forceinline__ void tile(int n, const DenseArgs& D, int slot, int r0, int c0, int lane, const void* Yin,
                                     void* Zout) {
    const int N = 1 << n;
    const int k = D.dense_k[slot];
    const int li = lane & 15, kq = lane >> 4;
    const c128* W = D.WT + (size_t)slot * N * N + c0 + li;  // B[kq][li] of k-step kb: W[(kb + kq) * N]
    const int row = r0 + li;
    const bool rv = row < D.SP;
    const cplx<Tg>* Arow = reinterpret_cast<const cplx<Tg>*>(Yin) + ((size_t)(rv ? row : 0) * D.L + k) * N + kq;
    const cplx<Tg> zero = mk((Tg)0, (Tg)0);
    v4d p1 = v4d{0.0, 0.0, 0.0, 0.0}, p2 = v4d{0.0, 0.0, 0.0, 0.0}, p3 = v4d{0.0, 0.0, 0.0, 0.0};
    cplx<Tg> a[U], an[U];
    c128 b[U], bn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        a[u] = rv ? Arow[4 * u] : zero;
        b[u] = W[(size_t)(4 * u + kq) * N];
    }
    for (int kb = 0; kb < N; kb += 4 * U) {
        const int kn = kb + 4 * U;
        if (kn < N) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                an[u] = rv ? Arow[kn + 4 * u] : zero;
                bn[u] = W[(size_t)(kn + 4 * u + kq) * N];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double ar = (double)a[u].x, ai = (double)a[u].y;
            p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b[u].x, p1, 0, 0, 0);
            p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, b[u].y, p2, 0, 0, 0);
            p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, b[u].x + b[u].y, p3, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[u] = an[u];
            b[u] = bn[u];
        }
    }
    cplx<Tg>* Z = reinterpret_cast<cplx<Tg>*>(Zout);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int sp = r0 + 4 * e + kq;
        if (sp < D.SP) Z[((size_t)sp * D.L + k) * N + c0 + li] = cast<Tg>(mk(p1[e] - p2[e], (p3[e] - p1[e]) - p2[e]));
    }
}

// Tiles of a pass: slot-major, then 16-output column tile, then 16-row tile, so consecutive tiles of a worker share
// W columns and a frequency's tiles sit on few workers.
__host__ __device__ inline int row_tiles(const DenseArgs& D) { return (D.SP + 15) / 16; }
__host__ __device__ inline long long tiles(int n, const DenseArgs& D) {
    return (long long)D.n_dense * ((1 << n) / 16) * row_tiles(D);
}

// Dense worker w of D.workers: its contiguous share of the pass's tiles, one after another.
template <typename Tg, int U>
__device__ __forceinline__ void worker(int n, const DenseArgs& D, int w, int lane, const void* Yin, void* Zout) {
    const long long T = tiles(n, D);
    const long long t0 = T * w / D.workers, t1 = T * (w + 1) / D.workers;
    const int gr = row_tiles(D), gc = (1 << n) / 16;
    for (long long t = t0; t < t1; ++t) {
        const int r = (int)(t % gr);
        const int c = (int)((t / gr) % gc);
        const int slot = (int)(t / ((long long)gr * gc));
        tile<Tg, U>(n, D, slot, 16 * r, 16 * c, lane, Yin, Zout);
    }
}

}  // namespace dense

}  // namespace adrt
