// Host side of the register row kernel (rows_reg.hpp): op-list decoding, default mapping, dispatch to the
// instantiation units rows_reg_{d,s,c}{1,2}.hip.
#include "rows_reg.hpp"

#include <atomic>

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

int dense_workers(int n, const DenseArgs& dn) {
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
    const long long T = dense::tiles(n, dn);
    return (int)(T < sm ? T : sm);
}

int launch_rows_reg(const RowsArgs& A, int prec, int loge, hipStream_t s) {
    RegArgs RA{};
    RA.r = A;
    if (!decode_ops(A.n, A.ops, A.nops, RA)) return E_INVALID;
    if (A.in_stride != A.N || A.out_stride != A.N) return E_INVALID;
    RA.group_by_k = A.group_by_k;
    RA.SP = A.SP;
    RA.korder = A.korder;
    RA.nk = A.nk > 0 ? A.nk : A.period;
    if (A.skip_dense && !A.group_by_k) return E_INVALID;  // only the frequency-grouped mapping can skip dense rows
    const bool lo1 = RA.lo == 1;
    if (prec == PREC_F64) return lo1 ? rows_reg_launch_d1(RA, loge, s) : rows_reg_launch_d2(RA, loge, s);
    if (A.comp) return lo1 ? rows_reg_launch_c1(RA, loge, s) : rows_reg_launch_c2(RA, loge, s);
    return lo1 ? rows_reg_launch_s1(RA, loge, s) : rows_reg_launch_s2(RA, loge, s);
}

}  // namespace adrt
