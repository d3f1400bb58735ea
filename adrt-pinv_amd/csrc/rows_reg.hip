// Host side of the register row kernel (rows_reg.hpp): op-list decoding, default mapping, dispatch to the
// instantiation units rows_reg_{d,s,c}{1,2}.hip and rows_reg_x2.hip.
#include "rows_reg.hpp"

#include <algorithm>
#include <atomic>
#include <vector>

namespace adrt {

// Decodes a reference op list into the kernel's macro-ops; returns false if the list needs the generic kernel.  The
// register kernel takes level 1 alone (B 1..1, BT 1..1) and ops over levels lo..n with one lo in {1, 2} shared by the
// whole list: a B lo..n directly followed by BT lo..n is the Gram sandwich, a B lo..n must otherwise end the list.
bool decode_ops(int n, const RowOp* ops, int nops, RegArgs& RA) {
    int m = 0, lo = 0;
    for (int i = 0; i < nops; ++i) {
        const RowOp& o = ops[i];
        const bool l1 = o.lo == 1 && o.hi == 1;
        if (o.kind != OP_P && l1) {
            RA.mop[m++] = o.kind == OP_B ? MOP_L1F : MOP_L1A;
            continue;
        }
        if (o.hi != n || (o.lo != 1 && o.lo != 2) || (lo != 0 && o.lo != lo)) return false;
        lo = o.lo;
        if (o.kind == OP_P) {
            RA.mop[m] = MOP_P;
        } else if (o.kind == OP_B) {
            if (i + 1 < nops && ops[i + 1].kind == OP_BT && ops[i + 1].lo == o.lo && ops[i + 1].hi == o.hi) {
                RA.mop[m] = MOP_GRAM;
                ++i;
            } else if (i + 1 == nops) {
                RA.mop[m] = MOP_BEND;
            } else {
                return false;
            }
        } else {
            RA.mop[m] = MOP_BTNAT;
        }
        ++m;
    }
    RA.nmops = m;
    RA.lo = lo != 0 ? lo : 1;
    return true;
}

// Default elements per lane (log2) of the register kernel for a size and precision; 0 = use the generic kernel.
int default_loge(int n, int prec) {
    if (n < 5 || n > 11) return 0;
    if (prec == PREC_F32) return 3;
    return 2;
}

int dense_workers(int n, int prec, const DenseArgs& dn) {
    static std::atomic<int> simds[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int sm = simds[dev].load(std::memory_order_relaxed);
    if (sm == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        sm = 4 * cus;
        simds[dev].store(sm, std::memory_order_relaxed);
    }
    const int per = dense_config_get(n, dense_class(prec, dn));
    if (per > 1) sm *= per;
    const long long T = dense::tiles(n, dn);
    return (int)(T < sm ? T : sm);
}

namespace {

// One thread per frequency; the recurrence is the one the register kernel evaluates without a table.
__global__ __launch_bounds__(256) void cheb_coef_kernel(int L, const int* deg, const double* a, const double* c,
                                                         const int* off, c128* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L) return;
    const double ak = a[k], ck = c[k];
    c128* o = out + off[k];
    o[0] = mk(0.0, 0.0);
    double om = 1.0;
    for (int it = 1; it < deg[k]; ++it) {
        const double omj = (it == 1) ? 1.0 / (1.0 - ck * ck / 2.0) : 1.0 / (1.0 - ck * ck * om / 4.0);
        o[it] = mk(omj * ak, omj - 1.0);
        om = omj;
    }
}

}  // namespace

namespace {
size_t cheb_table_head(int K) { return ((sizeof(int) * (size_t)K + 255) / 256) * 256; }

// One thread per (row k, power t): the same products the register kernel would form for its own table.
template <typename T>
__global__ __launch_bounds__(256) void ztab_kernel(int n, int K, const c128* z, cplx<T>* out) {
    const long long H = (1LL << n) / 2 + 1;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= (long long)K * H) return;
    const int k = (int)(i / H), t = (int)(i % H);
    c128 p2[11];
    ztab_powers(z[k], n, p2);
    out[i] = cast<T>(ztab_entry(p2, n, t));
}
}  // namespace

size_t ztab_bytes(int n, int K, int prec) {
    return (size_t)K * (((size_t)1 << n) / 2 + 1) * (prec == PREC_F64 ? sizeof(c128) : sizeof(c64));
}

int ztab_fill(int n, int K, const c128* d_z, int prec, void* d_tab, hipStream_t st) {
    ADRT_REQUIRE(n >= 1 && n <= 11 && K >= 1 && d_z && d_tab, "ztab_fill: bad argument");
    const long long cnt = (long long)K * ((1LL << n) / 2 + 1);
    const unsigned grid = (unsigned)((cnt + 255) / 256);
    if (prec == PREC_F64)
        hipLaunchKernelGGL(ztab_kernel<double>, dim3(grid), dim3(256), 0, st, n, K, d_z, (c128*)d_tab);
    else
        hipLaunchKernelGGL(ztab_kernel<float>, dim3(grid), dim3(256), 0, st, n, K, d_z, (c64*)d_tab);
    return launch_check("ztab_kernel");
}

size_t cheb_table_bytes(int K, const int* deg_host) {
    size_t e = 0;
    for (int k = 0; k < K; ++k) e += (size_t)std::max(deg_host[k], 1);
    return cheb_table_head(K) + sizeof(c128) * e;
}

int cheb_table_fill(int K, const int* deg_host, const int* d_deg, const double* d_a, const double* d_c, void* d_table,
                    hipStream_t st, const c128** wcoef, const int** wcoff) {
    ADRT_REQUIRE(K >= 1 && deg_host && d_deg && d_a && d_c && d_table, "cheb_table_fill: bad argument");
    std::vector<int> off(K);
    long long tot = 0;
    for (int k = 0; k < K; ++k) {
        off[k] = (int)tot;
        tot += std::max(deg_host[k], 1);
    }
    ADRT_REQUIRE(tot <= 0x7fffffffLL, "cheb_table_fill: Chebyshev degrees too large");
    int* d_off = (int*)d_table;
    c128* d_w = (c128*)((char*)d_table + cheb_table_head(K));
    ADRT_HIP(hipMemcpy(d_off, off.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(cheb_coef_kernel, dim3((K + 255) / 256), dim3(256), 0, st, K, d_deg, d_a, d_c, d_off, d_w);
    ADRT_TRY(launch_check("cheb_coef_kernel"));
    *wcoef = d_w;
    *wcoff = d_off;
    return OK;
}

int launch_rows_reg(const RowsArgs& A, int prec, int loge, hipStream_t s) {
    RegArgs RA{};
    RA.r = A;
    if (!decode_ops(A.n, A.ops, A.nops, RA)) return E_INVALID;
    for (int o = 0; o < RA.nmops; ++o)
        if (RA.mop[o] == MOP_P && !(A.wcoef && A.wcoff)) return E_INVALID;  // P needs the weight table
    if (A.in_stride != A.N || A.out_stride != A.N) return E_INVALID;
    RA.group_by_k = A.group_by_k;
    RA.SP = A.SP;
    RA.korder = A.korder;
    RA.nk = A.nk > 0 ? A.nk : A.period;
    if (A.skip_dense && !A.group_by_k) return E_INVALID;  // only the frequency-grouped mapping can skip dense rows
    const bool lo1 = RA.lo == 1;
    if (A.io64) {
        ADRT_REQUIRE(prec == PREC_F32 && A.comp && !lo1,
                     "rows: complex128 rows with a float core need the compensated FP32 kernel and levels 2..n");
        return rows_reg_launch_x2(RA, loge, s);
    }
    if (prec == PREC_F64) return lo1 ? rows_reg_launch_d1(RA, loge, s) : rows_reg_launch_d2(RA, loge, s);
    if (A.comp) return lo1 ? rows_reg_launch_c1(RA, loge, s) : rows_reg_launch_c2(RA, loge, s);
    return lo1 ? rows_reg_launch_s1(RA, loge, s) : rows_reg_launch_s2(RA, loge, s);
}

}  // namespace adrt
