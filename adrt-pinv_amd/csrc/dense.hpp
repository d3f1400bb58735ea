// One 16 x 16 output tile of the dense low-frequency GEMM on the FP64 matrix cores, shared by the stand-alone dense
// kernel (dense.hip) and the dense workers of the register row kernel (rows_reg.hpp).
//
// C[r][i] = sum_j Y[r][j] WT[j][i] for 16 rows r of one dense frequency and 16 outputs i, complex, as three real
// v_mfma_f64_16x16x4 products per k-step (Gauss: P1 += Yr Wr, P2 += Yi Wi, P3 += (Yr + Yi)(Wr + Wi), then
// C_re = P1 - P2, C_im = P3 - P1 - P2; the sums Yr + Yi of float rows are exact in double).  Frequencies k and L - k
// have conjugate pass matrices (conjugate twiddles, real coefficients in every op), so a unit of work may carry the
// rows of both: the second frequency's rows enter conjugated and leave conjugated, conj(conj(Y) W) = Y conj(W), and
// one stream of W serves the two.  Operand layouts of the
// 16x16x4 f64 instruction (checked on gfx950): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, and
// D[4 r + lane / 16][lane % 16] in accumulator element r.  The k-steps run in ascending order whatever U is, so
// every caller gets the same bits for a row.
#pragma once
#include "common.hpp"

namespace adrt {

// The dense rows of a pass in n_dense units: unit u multiplies rows sp * L + f (sp < SP) of frequency f =
// dense_k[u] & FREQ_MASK by W_s, s = wslot[u] (u when wslot is null; W^T stored, N x N), or by conj(W_s) when
// dense_k[u] has the CONJ bit, and, when pair_k is given and pair_k[u] >= 0, the rows of frequency pair_k[u] in
// the same way.  A unit has `rows` rows: SP, or 2 SP with pairs (0 means SP).  A frequency's rows get the same bits
// in a unit of their own as beside a partner.
struct DenseArgs {
    const int* dense_k = nullptr;
    const int* pair_k = nullptr;
    const int* wslot = nullptr;
    const c128* WT = nullptr;
    int n_dense = 0;
    int L = 0;
    int SP = 0;
    int rows = 0;
    int workers = 0;  // register row kernel: waves that take dense tiles (0 = none)
    int dblk = 0;     // register row kernel: leading blocks that hold those waves (a multiple of 8)
    int rt = 1, ct = 1;  // register row kernel: row and output tiles per work unit of a dense wave
};

namespace dense {

constexpr int CONJ = 1 << 30;
constexpr int FREQ_MASK = CONJ - 1;

typedef double v4d __attribute__((ext_vector_type(4)));

// Sign flip of x where the high word of `mask` has its top bit set (exact; one integer op).
__device__ __forceinline__ double flip(double x, unsigned mask) {
    return __hiloint2double(__double2hiint(x) ^ (int)mask, __double2loint(x));
}

// The block of RT x CT tiles (rows r0 .. r0 + 16 RT - 1, outputs c0 .. c0 + 16 CT - 1) of dense unit `slot`, computed
// by one wave (lane 0..63): each k-step loads RT row fragments and CT W fragments and runs 3 RT CT MFMAs, so a W
// fragment serves RT row tiles and a row fragment CT output tiles.  U k-steps (of 4) per load group, the next group's
// loads in flight while the current group's MFMAs run.  Every output element accumulates the same k-steps in the same
// order for any RT, CT and U, so the bits do not depend on the blocking.  Zout must not alias Yin (other waves still
// read the frequency's rows).
template <typename Tg, int U, int RT, int CT>
__device__ __forceinline__ void tile(int n, const DenseArgs& D, int slot, int r0, int c0, int lane, const void* Yin,
                                     void* Zout) {
    const int N = 1 << n;
    const int e1 = D.dense_k[slot], e2 = D.pair_k ? D.pair_k[slot] : -1;
    const int k = e1 & FREQ_MASK, kp = e2 < 0 ? -1 : (e2 & FREQ_MASK);
    const unsigned c1 = (e1 & CONJ) ? 0x80000000u : 0u, c2 = (e2 >= 0 && (e2 & CONJ)) ? 0x80000000u : 0u;
    const int ws = D.wslot ? D.wslot[slot] : slot;
    const int li = lane & 15, kq = lane >> 4;
    const c128* W = D.WT + (size_t)ws * N * N + c0 + li;  // B[kq][li] of k-step kb, tile ct: W[(kb + kq) * N + 16 ct]
    // A row past the unit's rows, or an output tile past N, only feeds outputs that are never stored (row i of the
    // product depends on row i of A alone, column j on column j of W alone), so such lanes and tiles read a valid
    // clamped address (row 0, tile 0) instead of selecting zeros: no predicated or flat loads in the k loop.  Rows
    // SP .. 2 SP - 1 of a paired unit are the partner's; a conjugated frequency's rows are read with their imaginary
    // parts negated (cj) and written conjugated.
    const cplx<Tg>* Arow[RT];
    unsigned cj[RT];
    bool cv[CT];
    int coff[CT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int row = r0 + 16 * rt + li;
        const bool second = row >= D.SP;
        const int sp = second ? row - D.SP : row;
        const bool partner = second && kp >= 0 && sp < D.SP;
        const int rr = (!second || partner) ? sp : 0;
        Arow[rt] = reinterpret_cast<const cplx<Tg>*>(Yin) + ((size_t)rr * D.L + (partner ? kp : k)) * N + kq;
        cj[rt] = partner ? c2 : c1;
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        cv[ct] = c0 + 16 * ct < N;
        coff[ct] = cv[ct] ? 16 * ct : 0;
    }
    v4d p1[RT][CT], p2[RT][CT], p3[RT][CT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) p1[rt][ct] = p2[rt][ct] = p3[rt][ct] = v4d{0.0, 0.0, 0.0, 0.0};
    cplx<Tg> a[U][RT], an[U][RT];
    c128 b[U][CT], bn[U][CT];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) a[u][rt] = Arow[rt][4 * u];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) b[u][ct] = W[(size_t)(4 * u + kq) * N + coff[ct]];
    }
    for (int kb = 0; kb < N; kb += 4 * U) {
        const int kn = kb + 4 * U;
        if (kn < N) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) an[u][rt] = Arow[rt][kn + 4 * u];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) bn[u][ct] = W[(size_t)(kn + 4 * u + kq) * N + coff[ct]];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double ar[RT], ai[RT], as[RT], bs[CT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                ar[rt] = (double)a[u][rt].x;
                ai[rt] = flip((double)a[u][rt].y, cj[rt]);
                as[rt] = ar[rt] + ai[rt];
            }
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) bs[ct] = b[u][ct].x + b[u][ct].y;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    p1[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt], b[u][ct].x, p1[rt][ct], 0, 0, 0);
                    p2[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[rt], b[u][ct].y, p2[rt][ct], 0, 0, 0);
                    p3[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[rt], bs[ct], p3[rt][ct], 0, 0, 0);
                }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) a[u][rt] = an[u][rt];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) b[u][ct] = bn[u][ct];
        }
    }
    cplx<Tg>* Z = reinterpret_cast<cplx<Tg>*>(Zout);
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = r0 + 16 * rt + 4 * e + kq;
                const bool second = row >= D.SP;
                const int sp = second ? row - D.SP : row;
                const v4d& q1 = p1[rt][ct];
                const v4d& q2 = p2[rt][ct];
                const v4d& q3 = p3[rt][ct];
                const double im = (q3[e] - q1[e]) - q2[e];
                if (cv[ct] && (!second || (kp >= 0 && sp < D.SP)))
                    Z[((size_t)sp * D.L + (second ? kp : k)) * N + c0 + 16 * ct + li] =
                        cast<Tg>(mk(q1[e] - q2[e], flip(im, second ? c2 : c1)));
            }
}

// Work units of a pass: blocks of RT x CT tiles, slot-major, then output block, then row block.
__host__ __device__ inline int unit_rows(const DenseArgs& D) { return D.rows > 0 ? D.rows : D.SP; }
__host__ __device__ inline int row_tiles(const DenseArgs& D) { return (unit_rows(D) + 15) / 16; }
__host__ __device__ inline long long units(int n, const DenseArgs& D, int RT, int CT) {
    return (long long)D.n_dense * (((1 << n) / 16 + CT - 1) / CT) * ((row_tiles(D) + RT - 1) / RT);
}

__host__ __device__ inline long long tiles(int n, const DenseArgs& D) { return units(n, D, D.rt, D.ct); }

// Dense worker w of D.workers: units w, w + workers, w + 2 workers, ... of the pass's units of RT x CT tiles, so the
// workers run consecutive units together: all output tiles of a frequency are in flight at once and share its rows
// through L2 (N=1024, 8 sinograms: 111.8 -> 107.1 ms per batch against contiguous shares of units).
template <typename Tg, int U, int RT, int CT>
__device__ __forceinline__ void worker_b(int n, const DenseArgs& D, int w, int lane, const void* Yin, void* Zout) {
    const int T = (int)units(n, D, RT, CT);
    const int gr = (row_tiles(D) + RT - 1) / RT, gc = ((1 << n) / 16 + CT - 1) / CT;
    for (int t = w; t < T; t += D.workers) {
        const int r = t % gr, c = (t / gr) % gc, slot = t / (gr * gc);
        tile<Tg, U, RT, CT>(n, D, slot, 16 * RT * r, 16 * CT * c, lane, Yin, Zout);
    }
}

// BR x BC: the blocking the calling kernel's register budget allows; the host sets (D.rt, D.ct) to it or to 1 x 1.
template <typename Tg, int U, int BR, int BC>
__device__ __forceinline__ void worker(int n, const DenseArgs& D, int w, int lane, const void* Yin, void* Zout) {
    if constexpr (BR * BC > 1) {
        if (D.rt == BR && D.ct == BC) {
            worker_b<Tg, U, BR, BC>(n, D, w, lane, Yin, Zout);
            return;
        }
    }
    worker_b<Tg, U, 1, 1>(n, D, w, lane, Yin, Zout);
}

}  // namespace dense

}  // namespace adrt
