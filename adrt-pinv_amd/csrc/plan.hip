// C ABI, part 2: device-resident inverse plans (L_p and hybrid), their operators and the whole apply
// (fastops.py:175-338 on the device), hipGraph capture, and the forward / transpose ADRT.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/adrt_hip.h"
#include "passes.hpp"
#include "rows.hpp"

namespace adrt {

// A caller's stream and device-memory budget (adrt_ctx_*): plans and buffers charge their bytes to it.
struct Context {
    std::mutex mu;
    hipStream_t stream = nullptr;
    unsigned long long budget = 0;  // 0: no limit
    unsigned long long used = 0;
    long long plans = 0;            // live plans created in the context
    std::unordered_map<void*, unsigned long long> bufs;  // live buffers and their sizes

    // reserves `bytes` against the budget, or fails with E_NOMEM and leaves the usage unchanged
    int charge(unsigned long long bytes, const char* what) {
        std::lock_guard<std::mutex> g(mu);
        if (budget != 0 && (bytes > budget || used > budget - bytes)) {
            set_error(std::string(what) + ": " + std::to_string(bytes) + " bytes exceed the context's budget (" +
                      std::to_string(budget - used) + " of " + std::to_string(budget) + " left)");
            return E_NOMEM;
        }
        used += bytes;
        return OK;
    }
    void release(unsigned long long bytes) {
        std::lock_guard<std::mutex> g(mu);
        used -= bytes < used ? bytes : used;
    }
};

struct Plan {
    adrt_plan_desc d{};
    int N = 0, L = 0;
    double tau = 0;
    // row parameters
    c128* z = nullptr;
    int* deg = nullptr;
    double* a = nullptr;
    double* c = nullptr;
    int* dslot = nullptr;
    c128* dFT = nullptr;
    void* ztab64 = nullptr;        // z_k^t, t <= N/2, per frequency, complex128 (ztab_fill)
    void* ztab32 = nullptr;        // the same in complex64 (FP32 plans)
    void* wtab = nullptr;          // Chebyshev step weights of every frequency (cheb_table_fill)
    const c128* wcoef = nullptr;   // (inside wtab)
    const int* wcoff = nullptr;
    int* korder = nullptr;  // frequencies by descending degree (block order of the row kernel)
    int* korder_nd = nullptr;  // the same without the dense frequencies
    int* dense_k = nullptr;    // dense frequencies (GEMM slots)
    // the GEMM's units (dense::CONJ marks a frequency applied with the conjugate of its partner's W): one frequency
    // each (unit_k, unit_w; n_units), or conjugate pairs sharing one W (pair_k1, pair_k2, pair_w; n_punits)
    int* unit_k = nullptr;
    int* unit_w = nullptr;
    int* pair_k1 = nullptr;
    int* pair_k2 = nullptr;
    int* pair_w = nullptr;
    int n_units = 0, n_punits = 0, n_pairs = 0;
    c128* WopT = nullptr;      // (tables, N, N) pass matrices, transposed
    c128* WbackT = nullptr;
    void* tw = nullptr;    // FFT twiddles exp(-2 pi i m / L), m < N, double
    void* tw32 = nullptr;  // the same rounded to float
    // scratch for max_batch sinograms
    void* Y = nullptr;   // (S, 2, L, N) complex rows
    void* Y2 = nullptr;  // (S, 2, L, N) output rows of the row pass (out of place: the dense GEMM reads Y whole)
    void* Sq = nullptr;  // (S, 4, N, N) real quadrant strips
    double* y = nullptr;
    double* xs = nullptr;
    double* res = nullptr;
    double* r = nullptr;
    double* g = nullptr;
    RowOp ops_back[MAX_OPS];
    int nback = 0;
    RowOp ops_op[MAX_OPS];
    int nop = 0;
    int comp = 0;  // compensated FP32 Chebyshev accumulation (desc precision ADRT_PREC_F32)
    // row-kernel launches stamp their start / end into these event pairs, in turn (timing; empty: off)
    std::vector<hipEvent_t> ev_rows0, ev_rows1;
    size_t ev_next = 0;
    bool borrowed = false;  // dFT / WopT / WbackT belong to the caller (desc dense_on_device == 2)
    Context* ctx = nullptr;        // the context the plan's bytes are charged to (adrt_plan_create_in), or null
    unsigned long long charged = 0;
};

namespace {

// blocks of the paired inverse FFT at a plan's max_batch from which its FP32 passes take it: two per CU.  With fewer
// its four column FFTs per block in turn can cost more latency than the halved traffic saves (whole applies, N=256:
// one sinogram 2.91 -> 3.50 ms with it, 32 sinograms 9.50 -> 9.56 ms, 63 at max_batch 64 18.0 -> 17.4 ms).
constexpr long long PAIR_IFFT_MIN_BLOCKS = 512;
// rows per frequency (2 per sinogram) up to which conjugate dense frequencies share one unit
constexpr int DENSE_PAIR_ROWS = 32;

size_t row_elem(int prec) { return prec == PREC_F64 ? sizeof(c128) : sizeof(c64); }
size_t strip_elem(int prec) { return prec == PREC_F64 ? sizeof(double) : sizeof(float); }

void free_plan(Plan* p) {
    if (!p) return;
    if (p->borrowed) p->dFT = p->WopT = p->WbackT = nullptr;
    void* bufs[] = {p->z, p->deg, p->a, p->c, p->wtab, p->ztab64, p->ztab32, p->dslot, p->dFT, p->korder, p->korder_nd, p->dense_k, p->unit_k, p->unit_w, p->pair_k1, p->pair_k2, p->pair_w, p->WopT, p->WbackT, p->tw, p->tw32, p->Y, p->Y2, p->Sq, p->y, p->xs, p->res, p->r, p->g};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

// The row ops on the plan's rows Y, written out of place to `out` (the dense GEMM's blocks read all rows of their
// frequency while others write).
// the next timing event pair of the plan (adrt_plan_time_rows / adrt_plan_time_rows_ring), if any
void timing_events(Plan* p, RowsArgs& A) {
    if (p->ev_rows0.empty()) return;
    const size_t i = p->ev_next++ % p->ev_rows0.size();
    A.ev_start = p->ev_rows0[i];
    A.ev_stop = p->ev_rows1[i];
}

int rows_pass(Plan* p, int prec, const RowOp* ops, int nops, int S, hipStream_t st, void* out,
              const c128* WT = nullptr, bool io64 = false) {
    RowsArgs A{};
    A.io64 = io64 ? 1 : 0;
    A.Y = p->Y;
    A.Z = out;
    A.rows = (long long)S * 2 * p->L;
    A.in_stride = A.out_stride = p->N;
    A.n = p->d.n;
    A.N = p->N;
    A.period = p->L;
    A.z = p->z;
    A.deg = p->deg;
    A.a = p->a;
    A.c = p->c;
    A.tau = p->tau;
    A.dslot = p->dslot;
    A.dFT = p->dFT;
    A.wcoef = p->wcoef;
    A.wcoff = p->wcoff;
    A.ztab = prec == PREC_F64 ? p->ztab64 : p->ztab32;
    A.nops = nops;
    for (int i = 0; i < nops; ++i) A.ops[i] = ops[i];
    A.group_by_k = 1;
    A.SP = 2 * S;
    A.cfg_class = (prec == PREC_F64 && p->d.precision == PREC_F32) ? 2 : prec;
    A.korder = p->korder;
    A.comp = prec == PREC_F32 ? p->comp : 0;
    if (WT && p->d.n_dense > 0) {
        // Chebyshev rows in the row kernel, dense rows as one GEMM with the pass matrix
        A.korder = p->korder_nd;
        A.nk = p->L - p->d.n_dense;
        A.skip_dense = 1;
        // Every pass from N = 16 on takes the FP64 MFMA tiles, one sinogram included: in extra waves of the register
        // row kernel's launch they stream W far faster than the vector GEMV of one block per frequency (one FP64
        // sinogram at N=256: 59 -> 25 us per pass, 3.90 -> 2.51 ms per inverse).  A tile's per-row arithmetic does
        // not depend on the batch size, so a sinogram gives the same bits alone and in a batch.  (In an FP32 plan a
        // last-bit difference of an FP64 pass would otherwise reach an FP32 rounding of the residual, whose
        // cancellation magnifies it, and grow to the FP32 level.)  Without the register row kernel the tiles are a
        // launch of their own.
        const bool mfma = p->d.n >= 4;
        int dense_done = 0;
        // conjugate pairs share one stream of W where W dominates (few rows per frequency); with many rows the
        // frequencies go alone (N=256, 128 sinograms: 495 against 509 us per row launch).  Same bits either way.
        const bool pairs = p->n_pairs > 0 && 2 * S <= DENSE_PAIR_ROWS;
        if (mfma && A.nk > 0) {
            A.dn.dense_k = pairs ? p->pair_k1 : p->unit_k;
            A.dn.pair_k = pairs ? p->pair_k2 : nullptr;
            A.dn.wslot = pairs ? p->pair_w : p->unit_w;
            A.dn.WT = WT;
            A.dn.n_dense = pairs ? p->n_punits : p->n_units;
            A.dn.L = p->L;
            A.dn.SP = 2 * S;
            A.dn.rows = pairs ? 4 * S : 2 * S;
        }
        if (A.nk > 0) {
            timing_events(p, A);
            ADRT_TRY(launch_rows(A, prec, st, &dense_done));
        }
        if (dense_done) return OK;
        // the dense GEMM's row type: complex128 when the row kernel took complex128 rows
        const int rp = io64 ? PREC_F64 : prec;
        if (mfma) {
            if (pairs)
                return launch_dense_mfma(p->d.n, p->L, 2 * S, p->n_punits, p->pair_k1, WT, p->Y, A.Z, rp, st,
                                         p->pair_k2, p->pair_w);
            return launch_dense_mfma(p->d.n, p->L, 2 * S, p->n_units, p->unit_k, WT, p->Y, A.Z, rp, st, nullptr,
                                     p->unit_w);
        }
        return launch_dense_gemm(p->d.n, p->L, 2 * S, p->d.n_dense, p->dense_k, WT, p->Y, A.Z, rp, st);
    }
    timing_events(p, A);
    return launch_rows(A, prec, st);
}

// One packed pass: strips -> FFT -> row ops -> iFFT -> unpack -> combine (fastops.py:209-220).
// mode 0: src is an image batch, mode 1: a sinogram batch.  If resid_y != null, out = resid_y - pass(src).
// l1_64 (FP32 passes): the FFTs, the combine and the explicit level 1 in double around the float rows.
int packed_pass(Plan* p, int prec, int mode, const double* src, const RowOp* ops, int nops, const double* resid_y,
                void* out, int S, hipStream_t st, bool out_f32 = false, bool l1_64 = false) {
    const int n = p->d.n;
    const int pf = l1_64 ? PREC_F64 : prec;  // precision of the FFTs and the combine
    const void* tw = pf == PREC_F64 ? p->tw : p->tw32;
    ADRT_TRY(launch_pack_fft(n, mode, src, p->Y, S, pf, tw, st));
    ADRT_TRY(rows_pass(p, prec, ops, nops, S, st, p->Y2, mode == 1 ? p->WbackT : p->WopT, l1_64));
    // FP32 passes of plans whose full batch gives the paired inverse two blocks per CU take it (two planes instead of
    // four strips); a plan decides by its max_batch, never by the batch of a call, so an image's bits do not depend on
    // how many others share the launch
    if ((long long)p->d.max_batch * ifft_pair_blocks(n, pf) >= PAIR_IFFT_MIN_BLOCKS) {
        ADRT_TRY(launch_ifft_pair(n, p->Y2, p->Sq, S, pf, tw, st));
        return launch_combine_pair(n, p->Sq, pf, resid_y, out, S, st, out_f32);
    }
    ADRT_TRY(launch_ifft_unpack(n, p->Y2, p->Sq, S, pf, tw, st));
    return launch_combine(n, p->Sq, pf, resid_y, out, S, st, out_f32);
}

// P (the outer preconditioner): g1_inv by fixed-degree Chebyshev on [4, 16], lorr, or the identity.
int precond(Plan* p, const double* v, double* out, int S, hipStream_t st) {
    const int n = p->d.n;
    const long long count = (long long)S * p->N * p->N;
    if (p->d.kind != ADRT_KIND_HYBRID || p->d.precond == ADRT_PRECOND_NONE) {
        if (out != v) ADRT_HIP(hipMemcpyAsync(out, v, sizeof(double) * count, hipMemcpyDeviceToDevice, st));
        return OK;
    }
    if (p->d.precond == ADRT_PRECOND_LORR) return launch_lorr(n, v, out, S, st);
    const int deg = p->d.g1_deg;
    if (deg - 1 <= 16)
        return launch_g1_fused(n, deg, 0, v, nullptr, nullptr, 1.0, 0.0, 0.0, out, S, st, PREC_F64, nullptr,
                               p->d.max_batch);
    const double lo = 4.0, hi = 16.0;
    const double c = (hi - lo) / (hi + lo), a = 2.0 / (hi + lo);
    // u_j lands in buffer (j % 2) with buffer (deg - 1) % 2 == out
    double* buf[2];
    buf[(deg - 1) % 2] = out;
    buf[deg % 2] = p->g;
    ADRT_TRY(launch_scale(v, a, buf[0], count, st));
    double om = 1.0;
    for (int k = 1; k < deg; ++k) {
        om = (k == 1) ? 1.0 / (1.0 - c * c / 2.0) : 1.0 / (1.0 - c * c * om / 4.0);
        const double* xk = buf[(k - 1) % 2];
        const double* xp = (k == 1) ? nullptr : buf[k % 2];
        ADRT_TRY(launch_g1_step(n, v, xk, xp, a, om, om - 1.0, buf[k % 2], S, st));
    }
    return OK;
}

int check_batch(Plan* p, int S) {
    ADRT_REQUIRE(p, "null plan");
    ADRT_REQUIRE(S >= 0 && S <= p->d.max_batch, "batch larger than the plan's max_batch");
    return OK;
}

// The dense GEMM's units, formed on the host from the descriptor (dense::CONJ marks a frequency applied with the
// conjugate of its partner's pass matrices): one frequency each (uk, uw) or conjugate pairs sharing one table
// (pk1, pk2, pw), and the number of tables the descriptor's pass matrices hold.
struct DenseUnits {
    std::vector<int> uk, uw, pk1, pk2, pw;
    int n_tables = 0;
};

int dense_units(const adrt_plan_desc* d, DenseUnits& U) {
    const int L = 2 << d->n, nd = d->n_dense;
    std::vector<int> slot_of(L, -1);
    U.n_tables = d->dense_w ? 0 : nd;
    for (int i = 0; i < nd; ++i) {
        const int k = d->dense_k[i];
        ADRT_REQUIRE(k >= 0 && k < L, "plan_create: dense_k out of range");
        ADRT_REQUIRE(slot_of[k] < 0, "plan_create: dense_k lists a frequency twice");
        slot_of[k] = i;
        if (d->dense_w) {
            ADRT_REQUIRE(d->dense_w[i] >= -1 && d->dense_w[i] < nd, "plan_create: dense_w out of range");
            U.n_tables = std::max(U.n_tables, d->dense_w[i] + 1);
        }
    }
    auto table = [&](int i) { return d->dense_w ? d->dense_w[i] : i; };
    // Conjugate partners: k and L - k when their twiddles are exact conjugates (z from the angles 2 pi fftfreq(2N)
    // always are).  Every op of a pass has real coefficients, so the partner's pass matrix is the conjugate of k's:
    // the partner's rows always run as conj(conj(y) W_k), alone or in one unit with k's rows, where one stream of W_k
    // serves both.  The partner's own table, if it has one, is never read.
    std::vector<char> used((size_t)nd, 0);
    for (int i = 0; i < nd; ++i) {
        if (used[i] || table(i) < 0) continue;
        used[i] = 1;
        const int k = d->dense_k[i], kp = (L - k) % L;
        const int j = slot_of[kp];
        const bool conj = j >= 0 && j != i && !used[j] && d->z[kp].re == d->z[k].re && d->z[kp].im == -d->z[k].im;
        U.uk.push_back(k);
        U.uw.push_back(table(i));
        U.pk1.push_back(k);
        U.pw.push_back(table(i));
        U.pk2.push_back(conj ? (kp | dense::CONJ) : -1);
        if (conj) {
            used[j] = 1;
            U.uk.push_back(kp | dense::CONJ);
            U.uw.push_back(table(i));
        }
    }
    for (int i = 0; i < nd; ++i)
        ADRT_REQUIRE(used[i], "plan_create: dense_w gives a frequency no pass matrices and no conjugate partner with them");
    return OK;
}

}  // namespace

}  // namespace adrt

using namespace adrt;

extern "C" {

unsigned long long adrt_plan_bytes(const adrt_plan_desc* d) {
    if (!d || d->n < 1 || d->n > 11) return 0;
    const size_t N = (size_t)1 << d->n, L = 2 * N, S = d->max_batch > 0 ? d->max_batch : 1;
    size_t b = 0;
    b += L * (sizeof(c128) + sizeof(int) + 2 * sizeof(double) + sizeof(int));
    b += ztab_bytes(d->n, (int)L, PREC_F64);  // twiddle tables
    if (d->precision != ADRT_PREC_F64) b += ztab_bytes(d->n, (int)L, PREC_F32);
    if (d->deg) {
        b += cheb_table_bytes((int)L, d->deg);  // Chebyshev weight table
    } else {
        const std::vector<int> zdeg(L, 0);
        b += cheb_table_bytes((int)L, zdeg.data());
    }
    if (d->dense_on_device != 2 && d->n_dense > 0) {
        size_t nt = (size_t)d->n_dense;  // pass-matrix tables (max(dense_w) + 1 with dense_w)
        if (d->dense_w && d->dense_k) {
            nt = 0;
            for (int i = 0; i < d->n_dense; ++i) nt = std::max(nt, (size_t)std::max(d->dense_w[i] + 1, 0));
        }
        b += ((d->dense_F ? (size_t)d->n_dense : 0) + ((d->dense_Wop && d->dense_Wback) ? 2 * nt : 0)) * N * N *
             sizeof(c128);
    }
    b += twiddle_bytes(d->n, PREC_F64) + twiddle_bytes(d->n, PREC_F32);
    b += 2 * S * 2 * L * N * row_elem(PREC_F64);
    b += S * 4 * N * N * strip_elem(PREC_F64);
    b += 5 * S * N * N * sizeof(double);
    return b;
}

int adrt_plan_create(adrt_plan** out, const adrt_plan_desc* d) {
    ADRT_REQUIRE(out && d, "plan_create: null pointer");
    *out = nullptr;
    ADRT_REQUIRE(d->kind == ADRT_KIND_LP || d->kind == ADRT_KIND_HYBRID || d->kind == ADRT_KIND_NORMAL,
                 "plan_create: kind must be LP, HYBRID or NORMAL");
    ADRT_REQUIRE(d->n >= 1 && d->n <= 11, "plan_create: need 1 <= n <= 11");
    ADRT_REQUIRE(d->precision >= ADRT_PREC_F64 && d->precision <= ADRT_PREC_F32_PLAIN,
                 "plan_create: precision must be 0, 1 or 2");
    ADRT_REQUIRE(d->precond >= 0 && d->precond <= 2, "plan_create: precond must be 0, 1 or 2");
    ADRT_REQUIRE(d->g1_deg >= 1, "plan_create: g1_deg must be >= 1");
    ADRT_REQUIRE(d->outer_deg >= 1, "plan_create: outer_deg must be >= 1");
    ADRT_REQUIRE(d->hi > 0 && d->lo > -d->hi, "plan_create: need hi > 0 and hi + lo > 0");
    ADRT_REQUIRE(d->max_batch >= 1, "plan_create: max_batch must be >= 1");
    ADRT_REQUIRE(d->refine >= 0, "plan_create: refine must be >= 0");
    ADRT_REQUIRE(d->dense_on_device >= 0 && d->dense_on_device <= 2, "plan_create: dense_on_device must be 0, 1 or 2");
    ADRT_REQUIRE(d->g1_fp64 == 0 || d->g1_fp64 == 1, "plan_create: g1_fp64 must be 0 or 1");
    ADRT_REQUIRE(d->l1_fp64 == 0 || d->l1_fp64 == 1, "plan_create: l1_fp64 must be 0 or 1");
    ADRT_REQUIRE(!d->l1_fp64 || (d->precision == ADRT_PREC_F32 && d->kind == ADRT_KIND_HYBRID && d->n >= 5),
                 "plan_create: l1_fp64 needs an FP32 (compensated) hybrid plan with n >= 5");
    ADRT_REQUIRE(d->z && ((d->deg && d->a && d->c) || d->kind == ADRT_KIND_NORMAL),
                 "plan_create: null row parameter array");
    ADRT_REQUIRE(d->kind != ADRT_KIND_NORMAL || d->n_dense == 0, "plan_create: the normal operator has no dense rows");
    ADRT_REQUIRE(d->n_dense >= 0 && d->n_dense <= (2 << d->n), "plan_create: bad n_dense");
    const bool has_w = d->dense_Wop && d->dense_Wback;
    ADRT_REQUIRE(d->n_dense == 0 || (d->dense_k && (d->dense_F || has_w)),
                 "plan_create: dense rows need dense_k and dense_F or the pass matrices");
    ADRT_REQUIRE(!d->dense_w || (has_w && d->n >= 4), "plan_create: dense_w needs the pass matrices and n >= 4");
    DenseUnits U;
    if (d->n_dense > 0) ADRT_TRY(dense_units(d, U));
    auto* p = new Plan();
    p->d = *d;
    p->comp = d->precision == ADRT_PREC_F32;
    if (d->precision == ADRT_PREC_F32_PLAIN) p->d.precision = PREC_F32;
    p->d.lam_max = nullptr;
    p->d.z = nullptr;
    p->d.deg = nullptr;
    p->d.a = p->d.c = nullptr;
    p->d.dense_k = nullptr;
    p->d.dense_F = nullptr;
    p->d.dense_w = nullptr;
    const int n = d->n, N = 1 << n, L = 2 * N;
    p->N = N;
    p->L = L;
    p->tau = d->tau;
    auto fail = [&](int st) {
        free_plan(p);
        return st;
    };
    auto alloc = [&](void** ptr, size_t bytes) -> int { return adrt_malloc(ptr, bytes); };
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
        if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("plan_create: upload failed");
            return E_HIP;
        }
        return OK;
    };
    int st;
    if ((st = alloc((void**)&p->z, sizeof(c128) * L)) || (st = up(p->z, d->z, sizeof(c128) * L))) return fail(st);
    // (the normal operator has no Chebyshev rows: its deg / a / c may be null and are zero)
    const std::vector<int> zdeg(d->deg ? 0 : L, 0);
    const std::vector<double> zac(d->a && d->c ? 0 : L, 0.0);
    const int* degs = d->deg ? d->deg : zdeg.data();
    const double* as = d->a && d->c ? d->a : zac.data();
    const double* cs = d->a && d->c ? d->c : zac.data();
    if ((st = alloc((void**)&p->deg, sizeof(int) * L)) || (st = up(p->deg, degs, sizeof(int) * L))) return fail(st);
    if ((st = alloc((void**)&p->a, sizeof(double) * L)) || (st = up(p->a, as, sizeof(double) * L))) return fail(st);
    if ((st = alloc((void**)&p->c, sizeof(double) * L)) || (st = up(p->c, cs, sizeof(double) * L))) return fail(st);
    if ((st = alloc(&p->wtab, cheb_table_bytes(L, degs))) ||
        (st = cheb_table_fill(L, degs, p->deg, p->a, p->c, p->wtab, nullptr, &p->wcoef, &p->wcoff)))
        return fail(st);
    if ((st = alloc(&p->ztab64, ztab_bytes(n, L, PREC_F64))) || (st = ztab_fill(n, L, p->z, PREC_F64, p->ztab64, nullptr)))
        return fail(st);
    if (p->d.precision == PREC_F32 &&
        ((st = alloc(&p->ztab32, ztab_bytes(n, L, PREC_F32))) || (st = ztab_fill(n, L, p->z, PREC_F32, p->ztab32, nullptr))))
        return fail(st);
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        set_error("plan_create: Chebyshev weight table failed");
        return fail(E_HIP);
    }
    if (d->n_dense > 0) {
        std::vector<int> slot(L, -1);
        for (int i = 0; i < d->n_dense; ++i) slot[d->dense_k[i]] = i;
        if ((st = alloc((void**)&p->dslot, sizeof(int) * L)) || (st = up(p->dslot, slot.data(), sizeof(int) * L)))
            return fail(st);
        const size_t tb = sizeof(c128) * (size_t)d->n_dense * N * N;
        if (d->dense_on_device == 2) {
            p->borrowed = true;
            p->dFT = (c128*)d->dense_F;
        } else if (d->dense_F && (st = alloc((void**)&p->dFT, tb))) {
            return fail(st);
        }
        if (d->dense_F && d->dense_on_device == 1) {
            if (hipMemcpy(p->dFT, d->dense_F, tb, hipMemcpyDeviceToDevice) != hipSuccess) {
                set_error("plan_create: device copy of the dense tables failed");
                return fail(E_HIP);
            }
        } else if (d->dense_F && d->dense_on_device == 0) {
            std::vector<c128> ft((size_t)d->n_dense * N * N);
            const c128* F = (const c128*)d->dense_F;
            for (int s = 0; s < d->n_dense; ++s)
                for (int i = 0; i < N; ++i)
                    for (int j = 0; j < N; ++j)
                        ft[(size_t)s * N * N + (size_t)j * N + i] = F[(size_t)s * N * N + (size_t)i * N + j];
            if ((st = up(p->dFT, ft.data(), tb))) return fail(st);
        }
    }
    // pass matrices of the dense rows (stored transposed for coalesced streaming): U.n_tables of them
    if (d->n_dense > 0 && d->dense_Wop && d->dense_Wback) {
        const size_t nn = (size_t)N * N, nt = (size_t)U.n_tables;
        std::vector<c128> t(d->dense_on_device ? 0 : nt * nn);
        auto transpose_up = [&](const adrt_c128* W, c128** dst) -> int {
            if (d->dense_on_device == 2) {
                *dst = (c128*)W;
                return OK;
            }
            if (d->dense_on_device) {
                int r = alloc((void**)dst, sizeof(c128) * nt * nn);
                if (r) return r;
                if (hipMemcpy(*dst, W, sizeof(c128) * nt * nn, hipMemcpyDeviceToDevice) != hipSuccess) {
                    set_error("plan_create: device copy of the pass matrices failed");
                    return E_HIP;
                }
                return OK;
            }
            const c128* w = (const c128*)W;
            for (size_t s = 0; s < nt; ++s)
                for (int i = 0; i < N; ++i)
                    for (int j = 0; j < N; ++j) t[s * nn + (size_t)j * N + i] = w[s * nn + (size_t)i * N + j];
            int r = alloc((void**)dst, sizeof(c128) * t.size());
            return r ? r : up(*dst, t.data(), sizeof(c128) * t.size());
        };
        if ((st = transpose_up(d->dense_Wop, &p->WopT)) || (st = transpose_up(d->dense_Wback, &p->WbackT)))
            return fail(st);
        if ((st = alloc((void**)&p->dense_k, sizeof(int) * d->n_dense)) ||
            (st = up(p->dense_k, d->dense_k, sizeof(int) * d->n_dense)))
            return fail(st);
        p->n_units = (int)U.uk.size();
        p->n_punits = (int)U.pk1.size();
        p->n_pairs = p->n_units - p->n_punits;
        const size_t ub = sizeof(int) * U.uk.size(), pb = sizeof(int) * U.pk1.size();
        if ((st = alloc((void**)&p->unit_k, ub)) || (st = up(p->unit_k, U.uk.data(), ub)) ||
            (st = alloc((void**)&p->unit_w, ub)) || (st = up(p->unit_w, U.uw.data(), ub)) ||
            (st = alloc((void**)&p->pair_k1, pb)) || (st = up(p->pair_k1, U.pk1.data(), pb)) ||
            (st = alloc((void**)&p->pair_k2, pb)) || (st = up(p->pair_k2, U.pk2.data(), pb)) ||
            (st = alloc((void**)&p->pair_w, pb)) || (st = up(p->pair_w, U.pw.data(), pb)))
            return fail(st);
    }
    p->d.dense_Wop = p->d.dense_Wback = nullptr;
    // block order of the row kernel: longest (dense first, then highest degree) frequencies first
    {
        std::vector<int> ord(L);
        std::vector<long long> key(L);
        for (int k = 0; k < L; ++k) {
            ord[k] = k;
            key[k] = (long long)degs[k];
        }
        for (int i = 0; i < d->n_dense; ++i) key[d->dense_k[i]] = 1LL << 40;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return key[x] > key[y]; });
        if ((st = alloc((void**)&p->korder, sizeof(int) * L)) || (st = up(p->korder, ord.data(), sizeof(int) * L)))
            return fail(st);
        // non-dense frequencies follow the dense ones in ord
        const int nd = d->n_dense;
        if (nd < L) {
            if ((st = alloc((void**)&p->korder_nd, sizeof(int) * (L - nd))) ||
                (st = up(p->korder_nd, ord.data() + nd, sizeof(int) * (L - nd))))
                return fail(st);
        }
    }
    // FFT twiddles exp(-2 pi i m / L) for m < N, from the exact angle in double
    {
        std::vector<c128> t64(N);
        std::vector<c64> t32(N);
        for (int m = 0; m < N; ++m) {
            const double ang = -2.0 * M_PI * (double)m / (double)L;
            t64[m] = mk(std::cos(ang), std::sin(ang));
            t32[m] = mk((float)t64[m].x, (float)t64[m].y);
        }
        if ((st = alloc(&p->tw, twiddle_bytes(n, PREC_F64))) || (st = up(p->tw, t64.data(), twiddle_bytes(n, PREC_F64))))
            return fail(st);
        if ((st = alloc(&p->tw32, twiddle_bytes(n, PREC_F32))) ||
            (st = up(p->tw32, t32.data(), twiddle_bytes(n, PREC_F32))))
            return fail(st);
    }
    const size_t S = d->max_batch, NN = (size_t)N * N;
    if ((st = alloc(&p->Y, S * 2 * L * N * row_elem(PREC_F64)))) return fail(st);
    if ((st = alloc(&p->Y2, S * 2 * L * N * row_elem(PREC_F64)))) return fail(st);
    if ((st = alloc(&p->Sq, S * 4 * NN * strip_elem(PREC_F64)))) return fail(st);
    if ((st = alloc((void**)&p->y, S * NN * sizeof(double)))) return fail(st);
    if ((st = alloc((void**)&p->xs, S * NN * sizeof(double)))) return fail(st);
    if ((st = alloc((void**)&p->res, S * NN * sizeof(double)))) return fail(st);
    if ((st = alloc((void**)&p->r, S * NN * sizeof(double)))) return fail(st);
    if ((st = alloc((void**)&p->g, S * NN * sizeof(double)))) return fail(st);
    // op lists (fastops.py:281, 285, 330, 335)
    if (d->kind == ADRT_KIND_HYBRID) {
        const RowOp back[3] = {{OP_BT, 2, n}, {OP_P, 2, n}, {OP_BT, 1, 1}};
        const RowOp gh[5] = {{OP_B, 1, 1}, {OP_P, 2, n}, {OP_B, 2, n}, {OP_BT, 2, n}, {OP_BT, 1, 1}};
        if (n < 2) {
            set_error("plan_create: the hybrid needs n >= 2");
            return fail(E_INVALID);
        }
        p->nback = 3;
        p->nop = 5;
        for (int i = 0; i < 3; ++i) p->ops_back[i] = back[i];
        for (int i = 0; i < 5; ++i) p->ops_op[i] = gh[i];
    } else if (d->kind == ADRT_KIND_NORMAL) {
        // R^T and R^T R (patches/adrt-gram-freq.patch:454-466)
        p->nback = 1;
        p->nop = 2;
        p->ops_back[0] = RowOp{OP_BT, 1, n};
        p->ops_op[0] = RowOp{OP_B, 1, n};
        p->ops_op[1] = RowOp{OP_BT, 1, n};
    } else {
        const RowOp back[2] = {{OP_BT, 1, n}, {OP_P, 1, n}};
        const RowOp fr[3] = {{OP_P, 1, n}, {OP_B, 1, n}, {OP_BT, 1, n}};
        p->nback = 2;
        p->nop = 3;
        for (int i = 0; i < 2; ++i) p->ops_back[i] = back[i];
        for (int i = 0; i < 3; ++i) p->ops_op[i] = fr[i];
    }
    *out = (adrt_plan*)p;
    return OK;
}

int adrt_plan_destroy(adrt_plan* plan) {
    auto* p = (Plan*)plan;
    if (p && p->ctx) {
        p->ctx->release(p->charged);
        std::lock_guard<std::mutex> g(p->ctx->mu);
        --p->ctx->plans;
    }
    free_plan(p);
    return OK;
}

int adrt_ctx_create(adrt_ctx** out, unsigned long long budget_bytes) {
    ADRT_REQUIRE(out, "ctx_create: null out pointer");
    *out = nullptr;
    auto* c = new Context();
    c->budget = budget_bytes;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("ctx_create: stream creation failed");
        return E_HIP;
    }
    *out = (adrt_ctx*)c;
    return OK;
}

int adrt_ctx_destroy(adrt_ctx* ctx) {
    auto* c = (Context*)ctx;
    if (!c) return OK;
    {
        std::lock_guard<std::mutex> g(c->mu);
        if (c->plans != 0 || !c->bufs.empty()) {
            set_error("ctx_destroy: " + std::to_string(c->plans) + " plans and " + std::to_string(c->bufs.size()) +
                      " buffers of the context are still alive");
            return E_STATE;
        }
    }
    ADRT_HIP(hipStreamSynchronize(c->stream));
    ADRT_HIP(hipStreamDestroy(c->stream));
    delete c;
    return OK;
}

void* adrt_ctx_stream(adrt_ctx* ctx) { return ctx ? (void*)((Context*)ctx)->stream : nullptr; }

int adrt_ctx_usage(adrt_ctx* ctx, unsigned long long* used_bytes, unsigned long long* budget_bytes) {
    auto* c = (Context*)ctx;
    ADRT_REQUIRE(c, "ctx_usage: null context");
    std::lock_guard<std::mutex> g(c->mu);
    if (used_bytes) *used_bytes = c->used;
    if (budget_bytes) *budget_bytes = c->budget;
    return OK;
}

int adrt_ctx_malloc(adrt_ctx* ctx, void** ptr, unsigned long long bytes) {
    auto* c = (Context*)ctx;
    ADRT_REQUIRE(c && ptr, "ctx_malloc: null context or pointer");
    *ptr = nullptr;
    ADRT_TRY(c->charge(bytes, "ctx_malloc"));
    const int st = adrt_malloc(ptr, bytes);
    if (st != OK) {
        c->release(bytes);
        return st;
    }
    std::lock_guard<std::mutex> g(c->mu);
    c->bufs[*ptr] = bytes;
    return OK;
}

int adrt_ctx_free(adrt_ctx* ctx, void* ptr) {
    auto* c = (Context*)ctx;
    ADRT_REQUIRE(c, "ctx_free: null context");
    if (!ptr) return OK;
    unsigned long long bytes = 0;
    {
        std::lock_guard<std::mutex> g(c->mu);
        auto it = c->bufs.find(ptr);
        ADRT_REQUIRE(it != c->bufs.end(), "ctx_free: the pointer is not a buffer of this context");
        bytes = it->second;
        c->bufs.erase(it);
    }
    c->release(bytes);
    return adrt_free(ptr);
}

int adrt_plan_create_in(adrt_ctx* ctx, adrt_plan** out, const adrt_plan_desc* desc) {
    auto* c = (Context*)ctx;
    ADRT_REQUIRE(c, "plan_create_in: null context");
    ADRT_REQUIRE(out && desc, "plan_create_in: null pointer");
    *out = nullptr;
    const unsigned long long bytes = adrt_plan_bytes(desc);
    if (bytes == 0) return adrt_plan_create(out, desc);  // a size out of range: plan_create names the error
    ADRT_TRY(c->charge(bytes, "plan_create_in"));
    const int st = adrt_plan_create(out, desc);
    if (st != OK) {
        c->release(bytes);
        return st;
    }
    auto* p = (Plan*)*out;
    p->ctx = c;
    p->charged = bytes;
    std::lock_guard<std::mutex> g(c->mu);
    ++c->plans;
    return OK;
}

int adrt_plan_set_outer(adrt_plan* plan, double lo, double hi, int outer_deg) {
    auto* p = (Plan*)plan;
    ADRT_REQUIRE(p, "plan_set_outer: null plan");
    ADRT_REQUIRE(hi > 0 && lo > -hi, "plan_set_outer: need hi > 0 and hi + lo > 0");
    ADRT_REQUIRE(outer_deg >= 1, "plan_set_outer: outer_deg must be >= 1");
    p->d.lo = lo;
    p->d.hi = hi;
    p->d.outer_deg = outer_deg;
    return OK;
}

int adrt_plan_backproject(adrt_plan* plan, const double* d_b, double* d_y, int S, void* stream) {
    auto* p = (Plan*)plan;
    ADRT_TRY(check_batch(p, S));
    if (S == 0) return OK;
    ADRT_REQUIRE(d_b && d_y, "plan_backproject: null buffer");
    return packed_pass(p, p->d.precision, 1, d_b, p->ops_back, p->nback, nullptr, d_y, S, (hipStream_t)stream);
}

int adrt_plan_operator(adrt_plan* plan, const double* d_x, double* d_y, int S, void* stream) {
    auto* p = (Plan*)plan;
    ADRT_TRY(check_batch(p, S));
    if (S == 0) return OK;
    ADRT_REQUIRE(d_x && d_y, "plan_operator: null buffer");
    return packed_pass(p, p->d.precision, 0, d_x, p->ops_op, p->nop, nullptr, d_y, S, (hipStream_t)stream);
}

int adrt_plan_precond(adrt_plan* plan, const double* d_x, double* d_y, int S, void* stream) {
    auto* p = (Plan*)plan;
    ADRT_TRY(check_batch(p, S));
    if (S == 0) return OK;
    ADRT_REQUIRE(d_x && d_y && d_x != d_y, "plan_precond: null or aliased buffer");
    return precond(p, d_x, d_y, S, (hipStream_t)stream);
}

// The row kernel alone on the plan's current rows (the last packed pass), out of place into a second buffer so that
// repeated launches see identical input: the dominant kernel of an apply, for timing with events.
int adrt_plan_rows_pass(adrt_plan* plan, int which, int S, void* stream) {
    auto* p = (Plan*)plan;
    ADRT_TRY(check_batch(p, S));
    ADRT_REQUIRE(which >= 0 && which <= 3, "plan_rows_pass: which must be 0 / 1 (backproject / operator ops) or 2 / 3 "
                                           "(the same in FP64)");
    if (S == 0) return OK;
    const int prec = which >= 2 ? PREC_F64 : p->d.precision;
    return (which % 2) == 0 ? rows_pass(p, prec, p->ops_back, p->nback, S, (hipStream_t)stream, p->Y2, p->WbackT)
                            : rows_pass(p, prec, p->ops_op, p->nop, S, (hipStream_t)stream, p->Y2, p->WopT);
}

int adrt_plan_time_rows(adrt_plan* plan, void* ev_begin, void* ev_end) {
    auto* p = (Plan*)plan;
    ADRT_REQUIRE(p, "plan_time_rows: null plan");
    ADRT_REQUIRE((ev_begin == nullptr) == (ev_end == nullptr), "plan_time_rows: give both events or neither");
    p->ev_rows0.clear();
    p->ev_rows1.clear();
    p->ev_next = 0;
    if (ev_begin) {
        p->ev_rows0.push_back((hipEvent_t)ev_begin);
        p->ev_rows1.push_back((hipEvent_t)ev_end);
    }
    return OK;
}

int adrt_plan_time_rows_ring(adrt_plan* plan, void* const* ev_begin, void* const* ev_end, int count) {
    auto* p = (Plan*)plan;
    ADRT_REQUIRE(p, "plan_time_rows_ring: null plan");
    ADRT_REQUIRE(count >= 0 && (count == 0 || (ev_begin && ev_end)), "plan_time_rows_ring: bad event arrays");
    for (int i = 0; i < count; ++i)
        ADRT_REQUIRE(ev_begin[i] && ev_end[i], "plan_time_rows_ring: null event");
    p->ev_rows0.assign((hipEvent_t const*)ev_begin, (hipEvent_t const*)ev_begin + count);
    p->ev_rows1.assign((hipEvent_t const*)ev_end, (hipEvent_t const*)ev_end + count);
    p->ev_next = 0;
    return OK;
}

int adrt_plan_time_rows_count(adrt_plan* plan, long long* stamped) {
    auto* p = (Plan*)plan;
    ADRT_REQUIRE(p && stamped, "plan_time_rows_count: null argument");
    *stamped = (long long)p->ev_next;
    return OK;
}

int adrt_plan_refine_steps(adrt_plan* plan) {
    auto* p = (Plan*)plan;
    if (!p) return 0;
    const int D = p->d.outer_deg;
    return p->d.precision == PREC_F32 ? (p->d.refine < D ? p->d.refine : D - 1) : 0;
}

// x = outer Chebyshev of degree D on P G with right-hand side y = backproject(b) (fastops.py:247-258, 287-288, 337-338)
int adrt_plan_apply(adrt_plan* plan, const double* d_b, double* d_x, int S, void* stream) {
    auto* p = (Plan*)plan;
    ADRT_TRY(check_batch(p, S));
    if (S == 0) return OK;
    ADRT_REQUIRE(d_b && d_x, "plan_apply: null buffer");
    hipStream_t st = (hipStream_t)stream;
    const long long count = (long long)S * p->N * p->N;
    const double c = (p->d.hi - p->d.lo) / (p->d.hi + p->d.lo), a = 2.0 / (p->d.hi + p->d.lo);
    const int D = p->d.outer_deg;
    // FP32 plans with refine > 0: the right-hand side and the last `refine` outer steps in FP64 (a mixed-precision
    // refinement: the FP64 steps pull the iterate to the FP64 operator's fixed point)
    const int prec = p->d.precision;
    const int refine = adrt_plan_refine_steps(plan);
    const int pb = refine > 0 ? PREC_F64 : prec;
    ADRT_TRY(packed_pass(p, pb, 1, d_b, p->ops_back, p->nback, nullptr, p->y, S, st, false,
                         p->d.l1_fp64 && pb == PREC_F32));
    // iterate x_j lands in buffer j % 2, with buffer (D - 1) % 2 == d_x
    double* X[2];
    X[(D - 1) % 2] = d_x;
    X[D % 2] = p->xs;
    const bool has_p = p->d.kind == ADRT_KIND_HYBRID && p->d.precond != ADRT_PRECOND_NONE;
    // g1_inv fused with the outer scale / update when its degree fits the fused kernel's halo
    const bool fused = has_p && p->d.precond == ADRT_PRECOND_G1 && p->d.g1_deg - 1 <= 16;
    const double* py = p->y;
    if (fused) {
        ADRT_TRY(launch_g1_fused(p->d.n, p->d.g1_deg, 1, p->y, nullptr, nullptr, a, 0.0, 0.0, X[0], S, st,
                                 (refine > 0 || p->d.g1_fp64) ? PREC_F64 : prec, nullptr, p->d.max_batch));
    } else {
        if (has_p) {
            ADRT_TRY(precond(p, p->y, p->r, S, st));
            py = p->r;
        }
        ADRT_TRY(launch_scale(py, a, X[0], count, st));
    }
    double om = 1.0;
    for (int k = 1; k < D; ++k) {
        om = (k == 1) ? 1.0 / (1.0 - c * c / 2.0) : 1.0 / (1.0 - c * c * om / 4.0);
        const double* xk = X[(k - 1) % 2];
        const double* xp = (k == 1) ? nullptr : X[k % 2];
        const int pk = (refine > 0 && k >= D - refine) ? PREC_F64 : prec;
        // the G_1 iteration of an FP32 step runs in float unless the plan asks for double (g1_fp64)
        const int pg = p->d.g1_fp64 ? PREC_F64 : pk;
        // an FP32 step's residual feeds only the float G_1 iteration, which rounds it to float: store it as float
        const bool res32 = fused && pg == PREC_F32;
        ADRT_TRY(packed_pass(p, pk, 0, xk, p->ops_op, p->nop, p->y, p->res, S, st, res32,
                             p->d.l1_fp64 && pk == PREC_F32));
        if (fused) {
            ADRT_TRY(launch_g1_fused(p->d.n, p->d.g1_deg, 2, res32 ? nullptr : p->res, xk, xp, a, om, om - 1.0,
                                     X[k % 2], S, st, pg, res32 ? reinterpret_cast<const float*>(p->res) : nullptr,
                                     p->d.max_batch));
            continue;
        }
        const double* rr = p->res;
        if (has_p) {
            ADRT_TRY(precond(p, p->res, p->r, S, st));
            rr = p->r;
        }
        ADRT_TRY(launch_cheb_update(xk, xp, rr, a, om, om - 1.0, X[k % 2], count, st));
    }
    return OK;
}

// One outer step of every lane of a continuous batch: adrt_plan_apply's loop turned inside out, so that sinograms
// can join a batch between steps.  Lanes are ordered oldest first, so the lanes of one kind of step are a contiguous
// run: FP64 refinement steps, then the plan's own steps, then the sinograms that start (step 0).  Each run is one
// pass over its lanes; the outer update takes each lane's own omega_k.  A lane's arithmetic is that of
// adrt_plan_apply step for step, so its result does not depend on which lanes share its steps.
int adrt_plan_tick(adrt_plan* plan, const int* k_lane, int S, const double* d_b, double* d_xcur, double* d_xnext,
                   void* stream) {
    auto* p = (Plan*)plan;
    ADRT_TRY(check_batch(p, S));
    if (S == 0) return OK;
    ADRT_REQUIRE(k_lane && d_xcur && d_xnext && d_xcur != d_xnext, "plan_tick: null or aliased buffer");
    ADRT_REQUIRE(S <= G1_MAX_LANES, "plan_tick: at most 64 lanes");
    const int D = p->d.outer_deg;
    for (int i = 0; i < S; ++i) {
        ADRT_REQUIRE(k_lane[i] >= 0 && k_lane[i] < D, "plan_tick: a lane's step must be in [0, outer_deg)");
        ADRT_REQUIRE(i == 0 || k_lane[i] <= k_lane[i - 1], "plan_tick: lanes must be ordered oldest first");
    }
    ADRT_REQUIRE(k_lane[S - 1] > 0 || d_b, "plan_tick: lanes at step 0 need their sinograms (d_b)");
    hipStream_t st = (hipStream_t)stream;
    const size_t NN = (size_t)p->N * p->N;
    const size_t BB = (size_t)4 * (2 * p->N - 1) * p->N;
    const double c = (p->d.hi - p->d.lo) / (p->d.hi + p->d.lo), a = 2.0 / (p->d.hi + p->d.lo);
    const int prec = p->d.precision;
    const int refine = adrt_plan_refine_steps(plan);
    const bool has_p = p->d.kind == ADRT_KIND_HYBRID && p->d.precond != ADRT_PRECOND_NONE;
    const bool fused = has_p && p->d.precond == ADRT_PRECOND_G1 && p->d.g1_deg - 1 <= 16;
    // omega_k by adrt_plan_apply's recurrence, so every lane's weights are the same doubles
    std::vector<double> omk((size_t)D, 1.0);
    double om = 1.0;
    for (int k = 1; k < D; ++k) {
        om = (k == 1) ? 1.0 / (1.0 - c * c / 2.0) : 1.0 / (1.0 - c * c * om / 4.0);
        omk[(size_t)k] = om;
    }
    auto step_prec = [&](int k) {
        if (k == 0) return refine > 0 ? PREC_F64 : prec;
        return (refine > 0 && k >= D - refine) ? PREC_F64 : prec;
    };
    for (int i = 0; i < S;) {
        const bool start = k_lane[i] == 0;
        const int pk = step_prec(k_lane[i]);
        int j = i + 1;
        while (j < S && (k_lane[j] == 0) == start && step_prec(k_lane[j]) == pk) ++j;
        const int G = j - i;
        double* y = p->y + i * NN;
        double* xn = d_xnext + i * NN;
        if (start) {
            ADRT_TRY(packed_pass(p, pk, 1, d_b + i * BB, p->ops_back, p->nback, nullptr, y, G, st, false,
                                 p->d.l1_fp64 && pk == PREC_F32));
            if (fused) {
                ADRT_TRY(launch_g1_fused(p->d.n, p->d.g1_deg, 1, y, nullptr, nullptr, a, 0.0, 0.0, xn, G, st,
                                         (refine > 0 || p->d.g1_fp64) ? PREC_F64 : prec, nullptr, p->d.max_batch));
            } else {
                const double* py = y;
                if (has_p) {
                    ADRT_TRY(precond(p, y, p->r, G, st));
                    py = p->r;
                }
                ADRT_TRY(launch_scale(py, a, xn, (long long)G * NN, st));
            }
            i = j;
            continue;
        }
        const double* xk = d_xcur + i * NN;
        const int pg = p->d.g1_fp64 ? PREC_F64 : pk;
        const bool res32 = fused && pg == PREC_F32;
        ADRT_TRY(packed_pass(p, pk, 0, xk, p->ops_op, p->nop, y, p->res, G, st, res32,
                             p->d.l1_fp64 && pk == PREC_F32));
        if (fused) {
            G1Lanes ln;
            for (int l = 0; l < G; ++l) {
                const int k = k_lane[i + l];
                ln.w1[l] = omk[(size_t)k];
                ln.w0[l] = omk[(size_t)k] - 1.0;
                if (k >= 2) ln.has_xp |= 1ull << l;
            }
            ADRT_TRY(launch_g1_fused(p->d.n, p->d.g1_deg, 2, res32 ? nullptr : p->res, xk, xn, a, 0.0, 0.0, xn, G,
                                     st, pg, res32 ? reinterpret_cast<const float*>(p->res) : nullptr,
                                     p->d.max_batch, &ln));
        } else {
            const double* rr = p->res;
            if (has_p) {
                ADRT_TRY(precond(p, p->res, p->r, G, st));
                rr = p->r;
            }
            for (int l = 0; l < G; ++l) {
                const int k = k_lane[i + l];
                ADRT_TRY(launch_cheb_update(xk + l * NN, k >= 2 ? xn + l * NN : nullptr, rr + l * NN, a,
                                            omk[(size_t)k], omk[(size_t)k] - 1.0, xn + l * NN, (long long)NN, st));
            }
        }
        i = j;
    }
    return OK;
}

// Moves a lane of a continuous batch: its right-hand side (kept in the plan) and its two iterates.
int adrt_plan_lane_move(adrt_plan* plan, int from, int to, double* d_xa, double* d_xb, void* stream) {
    auto* p = (Plan*)plan;
    ADRT_REQUIRE(p, "null plan");
    ADRT_REQUIRE(d_xa && d_xb, "plan_lane_move: null buffer");
    ADRT_REQUIRE(from >= 0 && to >= 0 && from < p->d.max_batch && to < p->d.max_batch,
                 "plan_lane_move: lanes must be in [0, max_batch)");
    if (from == to) return OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t NN = (size_t)p->N * p->N, bytes = NN * sizeof(double);
    ADRT_HIP(hipMemcpyAsync(p->y + to * NN, p->y + from * NN, bytes, hipMemcpyDeviceToDevice, st));
    ADRT_HIP(hipMemcpyAsync(d_xa + to * NN, d_xa + from * NN, bytes, hipMemcpyDeviceToDevice, st));
    ADRT_HIP(hipMemcpyAsync(d_xb + to * NN, d_xb + from * NN, bytes, hipMemcpyDeviceToDevice, st));
    return OK;
}

// ---------------------------------------------------------------- graphs
struct adrt_graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
};

int adrt_graph_begin(void* stream) {
    ADRT_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed));
    return OK;
}

int adrt_graph_end(void* stream, adrt_graph** out) {
    ADRT_REQUIRE(out, "graph_end: null out pointer");
    *out = nullptr;
    hipGraph_t g = nullptr;
    ADRT_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(err));
        return E_HIP;
    }
    auto* G = new adrt_graph();
    G->g = g;
    G->e = e;
    *out = G;
    return OK;
}

int adrt_graph_launch(adrt_graph* g, void* stream) {
    ADRT_REQUIRE(g && g->e, "graph_launch: null graph");
    ADRT_HIP(hipGraphLaunch(g->e, (hipStream_t)stream));
    return OK;
}

int adrt_graph_destroy(adrt_graph* g) {
    if (!g) return OK;
    if (g->e) (void)hipGraphExecDestroy(g->e);
    if (g->g) (void)hipGraphDestroy(g->g);
    delete g;
    return OK;
}

int adrt_stream_wait(void* waiting, void* signal) {
    hipEvent_t ev;
    ADRT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e1 = hipEventRecord(ev, (hipStream_t)signal);
    hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent((hipStream_t)waiting, ev, 0) : e1;
    (void)hipEventDestroy(ev);
    if (e2 != hipSuccess) {
        set_error(std::string("stream_wait failed: ") + hipGetErrorString(e2));
        return E_HIP;
    }
    return OK;
}

// ---------------------------------------------------------------- timing events (for in-process measurement)
int adrt_event_create(void** ev) {
    ADRT_REQUIRE(ev, "event_create: null pointer");
    hipEvent_t e;
    ADRT_HIP(hipEventCreate(&e));
    *ev = (void*)e;
    return OK;
}

int adrt_event_destroy(void* ev) {
    if (ev) ADRT_HIP(hipEventDestroy((hipEvent_t)ev));
    return OK;
}

int adrt_stream_wait_event(void* stream, void* ev) {
    ADRT_REQUIRE(ev, "stream_wait_event: null event");
    ADRT_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0));
    return OK;
}

int adrt_event_record(void* ev, void* stream) {
    ADRT_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return OK;
}

int adrt_event_query(void* ev) {
    const hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_error(std::string("hipEventQuery failed: ") + hipGetErrorString(e));
    return E_HIP;
}

int adrt_event_elapsed_ms(void* start, void* stop, float* ms) {
    ADRT_REQUIRE(ms, "event_elapsed: null pointer");
    ADRT_HIP(hipEventSynchronize((hipEvent_t)stop));
    ADRT_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return OK;
}

// ---------------------------------------------------------------- forward / transpose ADRT
unsigned long long adrt_work_bytes_fwd(int n, int dtype, int S) { return adrt_work_bytes(n, dtype, S, 0); }
unsigned long long adrt_work_bytes_T(int n, int dtype, int S) { return adrt_work_bytes(n, dtype, S, 1); }

int adrt_forward(int n, int dtype, const void* d_x, void* d_b, void* d_work, int S, void* stream) {
    return launch_adrt_forward(n, dtype, d_x, d_b, d_work, S, (hipStream_t)stream);
}

int adrt_transpose(int n, int dtype, const void* d_b, void* d_x, void* d_work, int S, void* stream) {
    return launch_adrt_transpose(n, dtype, d_b, d_x, d_work, S, (hipStream_t)stream);
}

}  // extern "C"
