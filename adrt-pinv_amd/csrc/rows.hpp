// Fused per-frequency row kernel: butterfly levels, adjoints and the Chebyshev filter of one row, in LDS.
//
// Index scheme.  The reference writes level m out of place (cpp/oblique.cpp:22-50): outputs l2w+2t+b from inputs
// l2w+t and l2w+w+t.  Here every level is applied IN PLACE on the pair of physical slots (p0, p0 + w) that differ
// in bit m-1.  After forward levels lo..hi the physical slot p holds logical index Lambda(p) = rotl_hi(...rotl_lo(p))
// (rotl_k rotates the low k bits left by one); for lo <= 2 this is the bit reversal of the low hi bits.  The twiddle
// of the pair at physical p0 on level m is t = low (m-1) bits of Lambda_{lo..m-1}(p0).  The adjoint runs the same
// pairs in reverse level order and returns the natural order, so a B..BT sandwich (the Gram M = B^H B inside the
// Chebyshev filter) never permutes data.
#pragma once
#include "dense.hpp"

namespace adrt {

struct RowsArgs {
    const void* Y;          // input rows, complex<Tg>, row r at Y + r * in_stride
    void* Z;                // output rows (may alias Y)
    long long rows;         // number of rows
    long long in_stride;    // complex elements between rows
    long long out_stride;
    int n, N;
    int period;             // parameter index of row r is r % period
    const c128* z;          // (period) z = exp(i w) of each parameter row
    const int* deg;         // (period) Chebyshev degree (P ops)
    const double* a;        // (period) Chebyshev constants
    const double* c;
    double tau;
    const int* dslot;       // (period) dense-table slot of the row or -1; null when there are no dense rows
    const c128* dFT;        // dense tables, slot s at dFT + s*N*N, stored transposed: dFT[j*N + i] = F[i][j]
    int nops;
    RowOp ops[MAX_OPS];
    // block mapping of the register kernel: group_by_k != 0 means rows = SP * period and a block owns rows of one
    // frequency (equal degree); korder (may be null) lists the frequencies in block order (descending degree).
    int group_by_k = 0;
    int SP = 1;
    const int* korder = nullptr;
    int nk = 0;          // frequencies listed in korder (0 = period)
    int skip_dense = 0;  // rows of dense frequencies are left untouched (a dense GEMM handles them)
    int loge = 0;  // register kernel elements per lane (log2); 0 = default for (n, precision), -1 = generic kernel
    int cfg_class = -1;  // rows_config class of the pass (-1: its precision; 2: FP64 refinement pass of an FP32 plan)
    // when set, the row kernel's launch stamps its own start and end into these events (hipExtLaunchKernel: the
    // dispatch's timestamps, not markers that may be processed while the previous kernel still runs)
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    int comp = 0;  // FP32 only: compensated (TwoSum) accumulation of the Chebyshev iterate
    // FP32 register kernel only: the rows in Y and Z are complex128, and a leading B 1..1 / trailing BT 1..1 of the
    // op list (the hybrid's explicit level) runs in double on them; the ops between run in float
    int io64 = 0;
    // Chebyshev step weights (w_j a, w_j - 1) of parameter row k at wcoef[wcoff[k] + j], 1 <= j < deg[k]
    // (cheb_table_fill); the register kernel takes op lists with P ops only when they are given
    const c128* wcoef = nullptr;
    const int* wcoff = nullptr;
    // z^t of parameter row k for t = 0 .. N/2 at ztab + k (N/2 + 1), in the compute type (ztab_fill); null: the
    // register kernel builds the powers itself
    const void* ztab = nullptr;
    // dense rows of the pass for the register kernel to compute as MFMA tiles beside the rows (n_dense = 0: none);
    // requires skip_dense and Z distinct from Y
    DenseArgs dn{};
};

// Host launcher (rows.hip).  prec selects the compute type of the butterflies and Chebyshev steps; the global row
// type is complex128 for PREC_F64 and complex64 for PREC_F32.  When args.dn names dense rows, *dense_done is set to 1
// if the register kernel computed them and 0 if the caller still has to (the generic kernel ignores args.dn).
int launch_rows(const RowsArgs& args, int prec, hipStream_t stream, int* dense_done = nullptr);

// Register-resident kernel (rows_reg.hip); E_INVALID if the op list or size needs the generic kernel.
int launch_rows_reg(const RowsArgs& args, int prec, int loge, hipStream_t stream);
// Dense waves the register kernel adds for args.dn: at most dense_config_get(n, prec) (default 1) per SIMD of the
// device, the row waves keeping the others, and never more than the pass has tiles.
int dense_workers(int n, int prec, const DenseArgs& dn);
int default_loge(int n, int prec);
// Table of the Chebyshev step weights of the register kernel's P op for K parameter rows, in one device allocation
// of cheb_table_bytes(K, deg) bytes: K offsets, then wcoef[off[k] + j] = (w_j a_k, w_j - 1) for 1 <= j < deg[k]
// (w_1 = 1 / (1 - c^2 / 2), w_j = 1 / (1 - c^2 w_(j-1) / 4), the reference's cpp/oblique.cpp:86-98), max(deg[k], 1)
// entries per row.  deg_host is the host copy of d_deg; the weights are computed on the stream, the offsets are
// copied before return.
// z^(2^b), b < logn, by repeated squaring, and z^t as the product of those of t's set bits (in double): the register
// kernel's twiddles, computed the same way by the kernel itself and by ztab_fill.
__device__ __forceinline__ void ztab_powers(c128 z, int logn, c128* p2) {
    c128 p = z;
    for (int j = 0; j < logn; ++j) {
        p2[j] = p;
        p = mul(p, p);
    }
}
__device__ __forceinline__ c128 ztab_entry(const c128* p2, int logn, int t) {
    c128 w = mk(1.0, 0.0);
    for (int b = 0; b < logn; ++b)
        if ((t >> b) & 1) w = mul(w, p2[b]);
    return w;
}
// Fills the (K, N/2 + 1) table of z_k^t in the precision's type (complex128 for PREC_F64, complex64 otherwise).
size_t ztab_bytes(int n, int K, int prec);
int ztab_fill(int n, int K, const c128* d_z, int prec, void* d_tab, hipStream_t st);

size_t cheb_table_bytes(int K, const int* deg_host);
int cheb_table_fill(int K, const int* deg_host, const int* d_deg, const double* d_a, const double* d_c, void* d_table,
                    hipStream_t st, const c128** wcoef, const int** wcoff);
int rows_config_get(int n, int prec);
int rows_config_set(int n, int prec, int loge);
// dense GEMM placement per (n, class): 0 default, -1 own launch, k > 0 at most k dense waves per SIMD in the row
// launch; class = precision (0 FP64, 1 FP32), plus 2 for passes of one sinogram (unset: as the precision's)
int dense_config_get(int n, int prec);
int dense_config_set(int n, int prec, int mode);
int dense_class(int prec, const DenseArgs& dn);

// Dense rows as a GEMM with the per-frequency pass matrices W^T (dense.hip).
int launch_dense_gemm(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                      int prec, hipStream_t st);

// The same on the FP64 matrix cores; with pair_k, n_dense units of conjugate frequency pairs (DenseArgs).
int launch_dense_mfma(int n, int L, int SP, int n_dense, const int* dense_k, const c128* WT, const void* Y, void* Z,
                      int prec, hipStream_t st, const int* pair_k = nullptr, const int* wslot = nullptr);

// Validates an op list against n; returns OK or E_INVALID with the error set.
int check_ops(int n, const RowOp* ops, int nops);

}  // namespace adrt
