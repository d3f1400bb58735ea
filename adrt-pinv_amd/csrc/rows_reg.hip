// Register-resident row kernel.
//
// A row of N = 2^n complex values is held by G = N / E lanes, E = 2^e values per lane, RPB = max(64, G) / G rows per
// block.  The n butterfly bits are split into passes of e consecutive bits ([0, e), [e, 2e), ...).  In pass K a lane
// holds the E values whose physical indices differ only in the pass's bits, so every level of the pass is a
// register-local radix-2 step (in-place scheme of rows.hpp: pair bit m-1, twiddle z^t with t = bitrev_{m-1} of the low
// bits).  Between passes the row goes through LDS once (store in one grouping, load in the next).  A Chebyshev step is
// forward passes 0..K, adjoint passes K..0, so the row leaves and re-enters pass 0's grouping, which is the natural
// order: the iterate z, its predecessor and y stay in registers across all steps.  Twiddles are precomputed per lane
// for every pass at row start (in double, then rounded to T).
//
// Rows of one block run the same macro-op program; P rows with different degrees iterate to the block's largest
// degree with the update predicated off, so every barrier is block-uniform.  The plan maps blocks to rows of one
// frequency (equal degree) in descending-degree order.
#include "rows.hpp"

namespace adrt {

// macro-ops decoded on the host from the reference op list (see decode_ops)
enum MacroOp : int { MOP_L1F = 0, MOP_L1A = 1, MOP_P = 2, MOP_GRAM = 3, MOP_BTNAT = 4, MOP_BEND = 5 };

struct RegArgs {
    RowsArgs r;
    int nmops;
    int mop[MAX_OPS], mlo[MAX_OPS], mhi[MAX_OPS];
    int group_by_k;      // blocks own rows of one frequency (rows = SP * period)
    int SP;              // rows per frequency when group_by_k
    const int* korder;   // frequency order of the blocks (may be null)
    int nk;              // frequencies scheduled (korder entries)
};

namespace {

__device__ __forceinline__ int brev_low(int x, int bits) {
    return bits <= 0 ? 0 : (int)(__brev((unsigned)x) >> (32 - bits));
}

// physical index held after forward levels 1..hi for natural index x, and vice versa (bitrev of the low hi bits)
__device__ __forceinline__ int lam(int x, int hi) {
    if (hi <= 1) return x;
    const int mask = (1 << hi) - 1;
    return (x & ~mask) | brev_low(x & mask, hi);
}

template <typename T, int LOGN, int LOGE>
struct Cfg {
    static constexpr int N = 1 << LOGN;
    static constexpr int E = 1 << LOGE;
    static constexpr int G = N / E;
    static constexpr int NT = G < 64 ? 64 : G;
    static constexpr int RPB = NT / G;
    static constexpr int NPASS = (LOGN + LOGE - 1) / LOGE;
    static constexpr int ROW = N + (N >> LOGE);  // padded row length in LDS
    __device__ static constexpr int b0(int K) { return K * LOGE; }
    __device__ static constexpr int kb(int K) { return (LOGN - K * LOGE) < LOGE ? (LOGN - K * LOGE) : LOGE; }
    __device__ static __forceinline__ int pad(int p) { return p + (p >> LOGE); }
    // physical index of lane-local element u in the grouping of pass K
    template <int K>
    __device__ static __forceinline__ int phys(int lane, int u) {
        constexpr int B0 = K * LOGE;
        constexpr int KB = kb(K);
        const int j = u & ((1 << KB) - 1);
        const int gs = u >> KB;
        const int q = lane + G * gs;
        return ((q >> B0) << (B0 + KB)) | (j << B0) | (q & ((1 << B0) - 1));
    }
};

template <typename T>
struct Row {
    cplx<T> z, zc;
};

// ---- one pass of levels, forward / adjoint, in registers.  In level i of a pass, the pairs of a lane that share a
// twiddle t (same group gs, same low bits jl) also share t z, computed once: the forward pair is then
// (f0 + t f1, f0 + (t z) f1) and the adjoint (g0 + g1, conj(t) g0 + conj(t z) g1), two FMAs per real component.
template <typename T, int LOGN, int LOGE, int K>
__device__ __forceinline__ void pass_fwd(cplx<T> (&v)[1 << LOGE], const cplx<T>* tw, cplx<T> z, int lo, int hi) {
    using C = Cfg<T, LOGN, LOGE>;
    constexpr int KB = C::kb(K);
    constexpr int B0 = K * LOGE;
    constexpr int PER = (1 << KB) - 1;
#pragma unroll
    for (int i = 0; i < KB; ++i) {
        const int m = B0 + i + 1;
        if (m < lo || m > hi) continue;
#pragma unroll
        for (int gs = 0; gs < (C::E >> KB); ++gs) {
#pragma unroll
            for (int jl = 0; jl < (1 << i); ++jl) {
                const bool one = B0 == 0 && jl == 0;  // t = z^0
                const cplx<T> t = one ? mk((T)1, (T)0) : tw[gs * PER + (1 << i) - 1 + jl];
                const cplx<T> tz = one ? z : mul(t, z);
#pragma unroll
                for (int h = 0; h < (1 << (KB - 1 - i)); ++h) {
                    const int u0 = (gs << KB) | (h << (i + 1)) | jl;
                    const int u1 = u0 | (1 << i);
                    const cplx<T> f0 = v[u0], f1 = v[u1];
                    v[u0] = one ? add(f0, f1) : cfma(t, f1, f0);
                    v[u1] = cfma(tz, f1, f0);
                }
            }
        }
    }
}

template <typename T, int LOGN, int LOGE, int K>
__device__ __forceinline__ void pass_adj(cplx<T> (&v)[1 << LOGE], const cplx<T>* tw, cplx<T> zc, int lo, int hi) {
    using C = Cfg<T, LOGN, LOGE>;
    constexpr int KB = C::kb(K);
    constexpr int B0 = K * LOGE;
    constexpr int PER = (1 << KB) - 1;
    const cplx<T> z = conj(zc);
#pragma unroll
    for (int i = KB - 1; i >= 0; --i) {
        const int m = B0 + i + 1;
        if (m < lo || m > hi) continue;
#pragma unroll
        for (int gs = 0; gs < (C::E >> KB); ++gs) {
#pragma unroll
            for (int jl = 0; jl < (1 << i); ++jl) {
                const bool one = B0 == 0 && jl == 0;
                const cplx<T> t = one ? mk((T)1, (T)0) : tw[gs * PER + (1 << i) - 1 + jl];
                const cplx<T> tz = one ? z : mul(t, z);
#pragma unroll
                for (int h = 0; h < (1 << (KB - 1 - i)); ++h) {
                    const int u0 = (gs << KB) | (h << (i + 1)) | jl;
                    const int u1 = u0 | (1 << i);
                    const cplx<T> g0 = v[u0], g1 = v[u1];
                    v[u0] = add(g0, g1);
                    v[u1] = one ? cfmac(tz, g1, g0) : cfmac(tz, g1, cmulc(t, g0));
                }
            }
        }
    }
}

template <typename T, int LOGN, int LOGE, int KA, int KB_>
__device__ __forceinline__ void transition(cplx<T> (&v)[1 << LOGE], cplx<T>* X, int lane) {
    using C = Cfg<T, LOGN, LOGE>;
#pragma unroll
    for (int u = 0; u < C::E; ++u) X[C::pad(C::template phys<KA>(lane, u))] = v[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < C::E; ++u) v[u] = X[C::pad(C::template phys<KB_>(lane, u))];
    __syncthreads();
}

// levels lo..hi forward then adjoint (the Gram B^H B), grouping 0 -> grouping 0
template <typename T, int LOGN, int LOGE, int K>
__device__ __forceinline__ void gram_rec(cplx<T> (&v)[1 << LOGE], cplx<T>* X, const cplx<T> (*tw)[1 << LOGE], int lane,
                                         cplx<T> z, cplx<T> zc, int lo, int hi) {
    using C = Cfg<T, LOGN, LOGE>;
    pass_fwd<T, LOGN, LOGE, K>(v, tw[K], z, lo, hi);
    if constexpr (K + 1 < C::NPASS) {
        if (hi > (K + 1) * LOGE) {
            transition<T, LOGN, LOGE, K, K + 1>(v, X, lane);
            gram_rec<T, LOGN, LOGE, K + 1>(v, X, tw, lane, z, zc, lo, hi);
            transition<T, LOGN, LOGE, K + 1, K>(v, X, lane);
        }
    }
    pass_adj<T, LOGN, LOGE, K>(v, tw[K], zc, lo, hi);
}

// adjoint passes Klast..0 starting from grouping Klast
template <typename T, int LOGN, int LOGE, int K>
__device__ __forceinline__ void adj_down(cplx<T> (&v)[1 << LOGE], cplx<T>* X, const cplx<T> (*tw)[1 << LOGE], int lane,
                                         cplx<T> zc, int lo, int hi) {
    pass_adj<T, LOGN, LOGE, K>(v, tw[K], zc, lo, hi);
    if constexpr (K > 0) {
        transition<T, LOGN, LOGE, K, K - 1>(v, X, lane);
        adj_down<T, LOGN, LOGE, K - 1>(v, X, tw, lane, zc, lo, hi);
    }
}

// adjoint levels lo..hi of natural-order data (grouping 0) -> natural order (grouping 0)
template <typename T, int LOGN, int LOGE, int K>
__device__ __forceinline__ void btnat_from(cplx<T> (&v)[1 << LOGE], cplx<T>* X, const cplx<T> (*tw)[1 << LOGE],
                                           int lane, cplx<T> zc, int lo, int hi, int klast) {
    using C = Cfg<T, LOGN, LOGE>;
    if (K == klast) {
#pragma unroll
        for (int u = 0; u < C::E; ++u) v[u] = X[C::pad(C::template phys<K>(lane, u))];
        __syncthreads();
        adj_down<T, LOGN, LOGE, K>(v, X, tw, lane, zc, lo, hi);
        return;
    }
    if constexpr (K + 1 < C::NPASS) btnat_from<T, LOGN, LOGE, K + 1>(v, X, tw, lane, zc, lo, hi, klast);
}

template <typename T, int LOGN, int LOGE, int K>
__device__ __forceinline__ void fwd_up(cplx<T> (&v)[1 << LOGE], cplx<T>* X, const cplx<T> (*tw)[1 << LOGE], int lane,
                                       cplx<T> z, int lo, int hi, int klast) {
    pass_fwd<T, LOGN, LOGE, K>(v, tw[K], z, lo, hi);
    if constexpr (K + 1 < Cfg<T, LOGN, LOGE>::NPASS) {
        if (K < klast) {
            transition<T, LOGN, LOGE, K, K + 1>(v, X, lane);
            fwd_up<T, LOGN, LOGE, K + 1>(v, X, tw, lane, z, lo, hi, klast);
            return;
        }
    }
    // store the physical row from grouping K (== klast)
#pragma unroll
    for (int u = 0; u < Cfg<T, LOGN, LOGE>::E; ++u) X[Cfg<T, LOGN, LOGE>::pad(Cfg<T, LOGN, LOGE>::template phys<K>(lane, u))] = v[u];
    __syncthreads();
}

template <typename T, int LOGN, int LOGE, bool COMP>
__global__ __launch_bounds__((Cfg<T, LOGN, LOGE>::NT)) void rows_reg_kernel(RegArgs RA) {
    using C = Cfg<T, LOGN, LOGE>;
    constexpr int E = C::E;
    const RowsArgs& A = RA.r;
    __shared__ cplx<T> Xs[C::RPB * C::ROW];
    __shared__ int sdeg[C::RPB];

    const int tid = threadIdx.x;
    const int rib = tid / C::G;   // row in block
    const int lane = tid % C::G;  // lane within the row
    cplx<T>* X = Xs + rib * C::ROW;

    // ---- which row
    long long row;
    bool valid;
    if (RA.group_by_k) {
        const int nspb = (RA.SP + C::RPB - 1) / C::RPB;
        const int kk = blockIdx.x / nspb;
        const int sg = blockIdx.x % nspb;
        const int k = RA.korder ? RA.korder[kk] : kk;
        const int sp = sg * C::RPB + rib;
        valid = sp < RA.SP;
        row = (long long)(valid ? sp : 0) * A.period + k;
    } else {
        row = (long long)blockIdx.x * C::RPB + rib;
        valid = row < A.rows;
        if (!valid) row = 0;
    }
    const int k = (int)(row % A.period);
    const c128 zk = A.z[k];
    const cplx<T> z = cast<T>(zk);
    const cplx<T> zc = conj(z);

    // ---- twiddles of every pass for this lane, from z^(2^j) in double
    cplx<T> tw[C::NPASS][E];
    {
        c128 p2[LOGN];
        c128 p = zk;
#pragma unroll
        for (int j = 0; j < LOGN; ++j) {
            p2[j] = p;
            p = mul(p, p);
        }
#pragma unroll
        for (int K = 0; K < C::NPASS; ++K) {
            const int B0 = K * LOGE;
            const int KB = (LOGN - B0) < LOGE ? (LOGN - B0) : LOGE;
            const int PER = (1 << KB) - 1;
#pragma unroll
            for (int gs = 0; gs < (E >> KB); ++gs) {
                const int q = lane + C::G * gs;
                const int qlow = q & ((1 << B0) - 1);
                const int rq = brev_low(qlow, B0);
#pragma unroll
                for (int i = 0; i < KB; ++i) {
#pragma unroll
                    for (int jl = 0; jl < (1 << i); ++jl) {
                        const int t = (rq << i) | brev_low(jl, i);
                        c128 w = mk(1.0, 0.0);
#pragma unroll
                        for (int b = 0; b < LOGN; ++b)
                            if ((t >> b) & 1) w = mul(w, p2[b]);
                        tw[K][gs * PER + (1 << i) - 1 + jl] = cast<T>(w);
                    }
                }
            }
        }
    }

    // ---- load the row in natural order (grouping 0: lane holds E consecutive values)
    cplx<T> v[E];
    const int x0 = lane * E;
    // (the global row type is T: rows are stored in the compute precision)
    const cplx<T>* Yr = reinterpret_cast<const cplx<T>*>(A.Y) + row * A.in_stride;
#pragma unroll
    for (int u = 0; u < E; ++u) v[u] = valid ? Yr[x0 + u] : mk((T)0, (T)0);

    for (int o = 0; o < RA.nmops; ++o) {
        const int op = RA.mop[o], lo = RA.mlo[o], hi = RA.mhi[o];
        if (op == MOP_L1F) {
            pass_fwd<T, LOGN, LOGE, 0>(v, tw[0], z, lo, hi);
        } else if (op == MOP_L1A) {
            pass_adj<T, LOGN, LOGE, 0>(v, tw[0], zc, lo, hi);
        } else if (op == MOP_GRAM) {
            gram_rec<T, LOGN, LOGE, 0>(v, X, tw, lane, z, zc, lo, hi);
        } else if (op == MOP_BTNAT) {
            // natural -> physical order of levels 1..hi through LDS, then the adjoint passes down to grouping 0
#pragma unroll
            for (int u = 0; u < E; ++u) X[C::pad(lam(x0 + u, hi))] = v[u];
            __syncthreads();
            const int klast = (hi - 1) / LOGE;
            btnat_from<T, LOGN, LOGE, 0>(v, X, tw, lane, zc, lo, hi, klast);
        } else if (op == MOP_BEND) {
            const int klast = (hi - 1) / LOGE;
            fwd_up<T, LOGN, LOGE, 0>(v, X, tw, lane, z, lo, hi, klast);
#pragma unroll
            for (int u = 0; u < E; ++u) v[u] = X[C::pad(lam(x0 + u, hi))];
            __syncthreads();
        } else {  // MOP_P
            const int slot = A.dslot ? A.dslot[k] : -1;
            const bool dense = valid && slot >= 0;
            const int d = (valid && !dense) ? A.deg[k] : 0;
            if (lane == 0) sdeg[rib] = d * 2 + (dense ? 1 : 0);
            __syncthreads();
            int dmax = 0, anyd = 0;
#pragma unroll
            for (int r = 0; r < C::RPB; ++r) {
                dmax = max(dmax, sdeg[r] >> 1);
                anyd |= sdeg[r] & 1;
            }
            __syncthreads();
            // the iterate is zi (+ zl when COMP: the rounding errors of its increments, kept by TwoSum)
            constexpr int EL = COMP ? E : 1;
            cplx<T> y[E], zi[E], zp[E], zl[EL], zpl[EL];
            const double ak = valid ? A.a[k] : 0.0, ck = valid ? A.c[k] : 0.0;
            const T aT = (T)ak, tau = (T)A.tau;
#pragma unroll
            for (int u = 0; u < E; ++u) {
                y[u] = v[u];
                zi[u] = scale(aT, y[u]);
                zp[u] = mk((T)0, (T)0);
                if constexpr (COMP) {
                    zl[u] = mk(fma(aT, y[u].x, -zi[u].x), fma(aT, y[u].y, -zi[u].y));
                    zpl[u] = mk((T)0, (T)0);
                }
            }
            double om = 1.0;
            for (int it = 1; it < dmax; ++it) {
                const double omj = (it == 1) ? 1.0 / (1.0 - ck * ck / 2.0) : 1.0 / (1.0 - ck * ck * om / 4.0);
#pragma unroll
                for (int u = 0; u < E; ++u) v[u] = zi[u];
                gram_rec<T, LOGN, LOGE, 0>(v, X, tw, lane, z, zc, lo, hi);
                if (it < d) {
                    const T w1 = (T)omj, w0 = (T)(omj - 1.0);
#pragma unroll
                    for (int u = 0; u < E; ++u) {
                        if constexpr (COMP) {
                            const cplx<T> r = sub(sub(sub(y[u], v[u]), scale(tau, zi[u])), scale(tau, zl[u]));
                            const cplx<T> dz = add(sub(zi[u], zp[u]), sub(zl[u], zpl[u]));
                            cplx<T> hi, err;
                            two_sum(zi[u], add(scale(w1, scale(aT, r)), scale(w0, dz)), hi, err);
                            zp[u] = zi[u];
                            zpl[u] = zl[u];
                            zi[u] = hi;
                            zl[u] = add(zl[u], err);
                        } else {
                            const cplx<T> r = sub(sub(y[u], v[u]), scale(tau, zi[u]));
                            const cplx<T> zn = add(add(zi[u], scale(w1, scale(aT, r))), scale(w0, sub(zi[u], zp[u])));
                            zp[u] = zi[u];
                            zi[u] = zn;
                        }
                    }
                    om = omj;
                }
            }
            if (anyd) {
                // exact Tikhonov filter z = F y in double for dense rows (F^T stored: coalesced over the lane's E)
#pragma unroll
                for (int u = 0; u < E; ++u) X[C::pad(x0 + u)] = y[u];
                __syncthreads();
                if (dense) {
                    const c128* F = A.dFT + (long long)slot * C::N * C::N;
                    c128 acc[E];
#pragma unroll
                    for (int u = 0; u < E; ++u) acc[u] = mk(0.0, 0.0);
                    for (int q = 0; q < C::N; ++q) {
                        const c128 yq = cast<double>(X[C::pad(q)]);
                        const c128* Fq = F + (long long)q * C::N + x0;
#pragma unroll
                        for (int u = 0; u < E; ++u) acc[u] = add(acc[u], mul(Fq[u], yq));
                    }
#pragma unroll
                    for (int u = 0; u < E; ++u) {
                        zi[u] = cast<T>(acc[u]);
                        if constexpr (COMP) zl[u] = cast<T>(sub(acc[u], cast<double>(zi[u])));
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int u = 0; u < E; ++u) {
                if constexpr (COMP)
                    v[u] = add(zi[u], zl[u]);
                else
                    v[u] = zi[u];
            }
        }
    }
    if (valid) {
        cplx<T>* Zr = reinterpret_cast<cplx<T>*>(A.Z) + row * A.out_stride;
#pragma unroll
        for (int u = 0; u < E; ++u) Zr[x0 + u] = v[u];
    }
}

template <typename T, int LOGN, int LOGE, bool COMP>
int launch_cfg(const RegArgs& RA, hipStream_t s) {
    using C = Cfg<T, LOGN, LOGE>;
    long long nblk;
    if (RA.group_by_k)
        nblk = (long long)RA.nk * ((RA.SP + C::RPB - 1) / C::RPB);
    else
        nblk = (RA.r.rows + C::RPB - 1) / C::RPB;
    if (nblk < 1) return OK;
    if (nblk > 0x7fffffffLL) {
        set_error("rows_reg: too many blocks");
        return E_INVALID;
    }
    hipLaunchKernelGGL((rows_reg_kernel<T, LOGN, LOGE, COMP>), dim3((unsigned)nblk), dim3(C::NT), 0, s, RA);
    return launch_check("rows_reg_kernel");
}

template <typename T, int LOGN, bool COMP>
int launch_n(const RegArgs& RA, int loge, hipStream_t s) {
    if constexpr (LOGN >= 2) {
        if (loge == 2) return launch_cfg<T, LOGN, 2, COMP>(RA, s);
    }
    if constexpr (LOGN >= 3) {
        if (loge == 3) return launch_cfg<T, LOGN, 3, COMP>(RA, s);
    }
    if constexpr (LOGN >= 4) {
        if (loge == 4) return launch_cfg<T, LOGN, 4, COMP>(RA, s);
    }
    set_error("rows_reg: unsupported elements-per-lane");
    return E_INVALID;
}

template <typename T, bool COMP>
int launch_t(const RegArgs& RA, int loge, hipStream_t s) {
    switch (RA.r.n) {
        case 5: return launch_n<T, 5, COMP>(RA, loge, s);
        case 6: return launch_n<T, 6, COMP>(RA, loge, s);
        case 7: return launch_n<T, 7, COMP>(RA, loge, s);
        case 8: return launch_n<T, 8, COMP>(RA, loge, s);
        case 9: return launch_n<T, 9, COMP>(RA, loge, s);
        case 10: return launch_n<T, 10, COMP>(RA, loge, s);
        case 11: return launch_n<T, 11, COMP>(RA, loge, s);
        default: set_error("rows_reg: n must be 5..11"); return E_INVALID;
    }
}

}  // namespace

// Decodes a reference op list into the kernel's macro-ops; returns false if the list needs the generic kernel.
bool decode_ops(const RowOp* ops, int nops, RegArgs& RA) {
    int m = 0;
    for (int i = 0; i < nops; ++i) {
        const RowOp& o = ops[i];
        if (o.lo > 2) return false;
        if (o.kind == OP_P) {
            RA.mop[m] = MOP_P;
        } else if (o.kind == OP_B) {
            if (o.hi == 1) {
                RA.mop[m] = MOP_L1F;
            } else if (i + 1 < nops && ops[i + 1].kind == OP_BT && ops[i + 1].lo == o.lo && ops[i + 1].hi == o.hi) {
                RA.mop[m] = MOP_GRAM;
                ++i;
            } else if (i + 1 == nops) {
                RA.mop[m] = MOP_BEND;
            } else {
                return false;
            }
        } else {
            RA.mop[m] = o.hi == 1 ? MOP_L1A : MOP_BTNAT;
        }
        RA.mlo[m] = o.lo;
        RA.mhi[m] = o.hi;
        ++m;
    }
    RA.nmops = m;
    return true;
}

// Default elements per lane (log2) of the register kernel for a size and precision; 0 = use the generic kernel.
int default_loge(int n, int prec) {
    if (n < 5 || n > 11) return 0;
    if (prec == PREC_F32) return 3;
    return 2;
}

int launch_rows_reg(const RowsArgs& A, int prec, int loge, hipStream_t s) {
    RegArgs RA{};
    RA.r = A;
    if (!decode_ops(A.ops, A.nops, RA)) return E_INVALID;
    if (A.in_stride != A.N || A.out_stride != A.N) return E_INVALID;
    RA.group_by_k = A.group_by_k;
    RA.SP = A.SP;
    RA.korder = A.korder;
    RA.nk = A.nk > 0 ? A.nk : A.period;
    if (A.skip_dense && !A.group_by_k) return E_INVALID;  // only the frequency-grouped mapping can skip dense rows
    if (prec == PREC_F64) return launch_t<double, false>(RA, loge, s);
    if (A.comp) return launch_t<float, true>(RA, loge, s);
    return launch_t<float, false>(RA, loge, s);
}

}  // namespace adrt
