// Register-resident row kernel (templates; instantiated per compute type and lower level by rows_reg_*.hip).
//
// A row of N = 2^n complex values is held by G = N / E lanes, E = 2^e values per lane, RPB = max(64, G) / G rows per
// block.  The n butterfly bits are split into passes of e consecutive bits ([0, e), [e, 2e), ...).  In pass K a lane
// holds the E values whose physical indices differ only in the pass's bits, so every level of the pass is a
// register-local radix-2 step (in-place scheme of rows.hpp: pair bit m-1, twiddle z^t with t = bitrev_{m-1} of the low
// bits).  Between passes the row goes through LDS once (store in one grouping, load in the next).  A Chebyshev step is
// forward passes 0..K, adjoint passes K..0, so the row leaves and re-enters pass 0's grouping, which is the natural
// order: the iterate z, its predecessor and y stay in registers across all steps.
//
// The level range of every op is a compile-time constant: levels LO..n for the Chebyshev filter, the Gram sandwich
// and the end-of-list forward / adjoint ops (LO = 1 for L_p, 2 for the hybrid), level 1 alone for the explicit level.
// Those are the only ranges the inverses use (fastops.py:281-285, 330-335), so the passes are straight-line code.
// Each lane keeps the pair (z^t, z^(t+1)) of every twiddle slot it uses in registers: the forward pair is
// (f0 + z^t f1, f0 + z^(t+1) f1), the adjoint (g0 + g1, conj(z^t) g0 + conj(z^(t+1)) g1), two FMAs per real component,
// and no twiddle product is recomputed inside the Chebyshev loop.  The twiddles come from a table of z^t, t <= N/2, in
// the row's LDS at start: copied from the plan's per-frequency table (ztab_fill), or built by the row's lanes
// (double products of z^(2^b), rounded to the compute type) when the caller has none.  The innermost level n runs
// forward and adjoint as one halved step on h = (z^t + z^(t+1)) / 2 (inner_half), and the Chebyshev step weights come
// from the plan's table (cheb_table_fill).
//
// Rows of one block run the same macro-op program; P rows with different degrees iterate to the block's largest
// degree with the update predicated off, so every barrier is block-uniform.  The plan maps blocks to rows of one
// frequency (equal degree) in descending-degree order.
//
// A pass with dense rows (RowsArgs::dn) puts its dense GEMM tiles in the same launch: the leading blocks are dense
// workers, at most 1, 2 or 4 waves per SIMD (dense_config_get, chosen by the autotuner), each taking a contiguous
// share of the 16 x 16 MFMA tiles (dense.hpp), while the row blocks behind them fill the remaining wave slots.  The tiles' loads and matrix-core work then overlap the
// rows' vector work instead of running as a launch of their own after it.
#pragma once
#include <type_traits>

#include "rows.hpp"

namespace adrt {

// macro-ops decoded on the host from the reference op list (rows_reg.hip: decode_ops)
enum MacroOp : int { MOP_L1F = 0, MOP_L1A = 1, MOP_P = 2, MOP_GRAM = 3, MOP_BTNAT = 4, MOP_BEND = 5 };

struct RegArgs {
    RowsArgs r;
    int nmops;
    int mop[MAX_OPS];
    int lo;              // lower level of every P / GRAM / BTNAT / BEND macro-op (1 or 2); their upper level is n
    int group_by_k;      // blocks own rows of one frequency (rows = SP * period)
    int SP;              // rows per frequency when group_by_k
    const int* korder;   // frequency order of the blocks (may be null)
    int nk;              // frequencies scheduled (korder entries)
    int iter = 1;        // group_by_k: row groups (of RPB rows) a block runs one after another (set by the launcher)
};

// Launchers of one compute type and lower level, rows_reg_<type><lo>.hip (d: double, s: float, c: float with the
// TwoSum-compensated Chebyshev iterate); E_INVALID for an unsupported size.
int rows_reg_launch_d1(const RegArgs& RA, int loge, hipStream_t s);
int rows_reg_launch_d2(const RegArgs& RA, int loge, hipStream_t s);
int rows_reg_launch_s1(const RegArgs& RA, int loge, hipStream_t s);
int rows_reg_launch_s2(const RegArgs& RA, int loge, hipStream_t s);
int rows_reg_launch_c1(const RegArgs& RA, int loge, hipStream_t s);
int rows_reg_launch_c2(const RegArgs& RA, int loge, hipStream_t s);
// float with TwoSum, levels 2..n, complex128 rows in and out with the explicit level 1 in double (RowsArgs::io64)
int rows_reg_launch_x2(const RegArgs& RA, int loge, hipStream_t s);

namespace rreg {

__device__ __forceinline__ int brev_low(int x, int bits) {
    return bits <= 0 ? 0 : (int)(__brev((unsigned)x) >> (32 - bits));
}

// physical index held after forward levels 1..hi for natural index x, and vice versa (bitrev of the low hi bits)
__device__ __forceinline__ int lam(int x, int hi) {
    if (hi <= 1) return x;
    const int mask = (1 << hi) - 1;
    return (x & ~mask) | brev_low(x & mask, hi);
}

// LDS layout of the pass exchanges.  A row value at physical index p sits at LDS entry addr(p) = sum of w[b] over
// the set bits b of p: linear in the bits, so the E values a lane exchanges sit at one per-lane base plus
// compile-time offsets (immediate offsets of ds_read / ds_write).  In pass K a lane holds the E values whose
// physical indices differ in the pass's element bits [K e, K e + e); the lane's own index bits fill the other
// positions in the order ord[K]: bit i of the extended lane index q = lane + G * group (the group of the element in a
// narrower last pass, else 0) goes to physical bit ord[K][i].
//
// Default: natural order and w[b] = 2^b + (2^b >> e), i.e. addr(p) = p + (p >> e).  The tuned configurations below
// (every pass e bits wide) use weights whose 2-adic valuations make every 16-lane group of every grouping reach 16
// distinct 8-byte bank pairs (addr mod 16), the lane bits 0..3 of each pass being the positions of valuation 0, 1,
// 2, 3: the 128-byte-per-clock LDS then serves each 64-lane b64 access in four conflict-free cycles, where the
// default conflicts twice in the last two groupings.  (The default padding has no such solution: three groupings
// share lane bits pairwise.)  Pass 0 keeps its first six lane bits on physical bits e..e+5, so a wavefront's row
// load stays one contiguous span, and the passes that exchange across wavefronts are the same ones as in natural
// order.  (N = 256 and N = 512 with 8 values per lane have such layouts too, but measured 3-4% slower than the
// default in alternating runs on one box, so they keep the default; N = 1024 with 4 values per lane gained 2-4%.)
template <int LOGN, int LOGE>
struct Layout {
    static constexpr bool custom = false;
    static constexpr int w[1] = {0};
    static constexpr int ord[1][1] = {{0}};
};
template <>
struct Layout<8, 2> {  // N = 256, 4 values per lane
    static constexpr bool custom = true;
    static constexpr int w[8] = {4, 2, 34, 8, 68, 1, 136, 17};
    static constexpr int ord[4][6] = {{5, 2, 4, 3, 6, 7}, {5, 1, 0, 6, 4, 7}, {7, 1, 0, 3, 2, 6}, {5, 1, 0, 3, 2, 4}};
};
template <>
struct Layout<10, 2> {  // N = 1024, 4 values per lane
    static constexpr bool custom = true;
    static constexpr int w[10] = {8, 1, 2, 4, 136, 68, 34, 17, 289, 578};
    static constexpr int ord[5][8] = {{7, 2, 3, 4, 5, 6, 8, 9}, {1, 6, 5, 0, 4, 7, 8, 9}, {1, 2, 3, 0, 6, 7, 8, 9},
                                      {1, 2, 3, 0, 4, 5, 8, 9}, {1, 2, 3, 0, 4, 5, 6, 7}};
};
template <>
struct Layout<10, 3> {  // N = 1024, 8 values per lane (passes of 3, 3, 3, 1 bits)
    static constexpr bool custom = true;
    static constexpr int w[10] = {2, 8, 1, 4, 136, 17, 34, 578, 68, 289};
    static constexpr int ord[4][9] = {{5, 6, 3, 4, 7, 8, 9, 0, 0}, {2, 0, 8, 1, 6, 7, 9, 0, 0},
                                      {2, 0, 3, 1, 4, 5, 9, 0, 0}, {2, 0, 3, 1, 4, 5, 6, 7, 8}};
};

template <typename T, int LOGN, int LOGE>
struct Cfg {
    static constexpr int N = 1 << LOGN;
    static constexpr int E = 1 << LOGE;
    static constexpr int G = N / E;
    static constexpr int NT = G < 64 ? 64 : G;
    static constexpr int RPB = NT / G;
    static constexpr int NPASS = (LOGN + LOGE - 1) / LOGE;
    static constexpr int ROW = N + (N >> LOGE);  // padded row length in LDS (also holds the z^t table, N/2 + 1)
    // Waves per SIMD the register allocation must allow (measured on MI355X): float rows with E = 4 from N = 128 on
    // run faster at 4 waves even with a few spilled registers than at their natural 3, and E = 8 up to N = 512 and at
    // N = 2048 at 3 rather than 2 (N = 2048, 4 sinograms: 340 -> 311 ms per batch).  At N = 1024 the extra wave makes
    // E = 8 about as fast as E = 4 with 8 sinograms (78.6 -> 73.7 against 74.9 ms) but slower with one (32.0 -> 37.8
    // against 27.0 ms); kept at 2 there, the tuner keeps E = 4, which serves single requests fastest.
    static constexpr int MINW = (sizeof(T) == 4 && LOGE == 2 && LOGN >= 7)                    ? 4
                                : (sizeof(T) == 4 && LOGE == 3 && (LOGN <= 9 || LOGN == 11)) ? 3
                                                                                              : 1;
    // Tiles per unit of the dense waves in the same launch (rows x outputs): 1 x 2 (each row fragment serves both
    // output tiles) where the rows leave the registers for it (3 waves per SIMD); at 4 waves the blocked tile would
    // spill, and the kernels without an occupancy bound would lose a wave to it.
    static constexpr int DENSE_BR = 1, DENSE_BC = MINW == 3 ? 2 : 1;
    __host__ __device__ static constexpr int kb(int K) { return (LOGN - K * LOGE) < LOGE ? (LOGN - K * LOGE) : LOGE; }
    using LY = Layout<LOGN, LOGE>;
    // LDS entry of physical index p (linear in p's bits)
    __host__ __device__ static constexpr int addr(int p) {
        if constexpr (LY::custom) {
            int a = 0;
            for (int b = 0; b < LOGN; ++b)
                if ((p >> b) & 1) a += LY::w[b];
            return a;
        } else {
            return p + (p >> LOGE);
        }
    }
    // physical index of lane-local element u in the grouping of pass K (lane + G * (u >> KB) is the lane's index
    // extended by the element's group in a narrower last pass), and its inverse: the lane holding p in grouping K
    __host__ __device__ static constexpr int phys_c(int K, int lane, int u) {
        const int B0 = K * LOGE, KB = kb(K);
        const int q = lane + G * (u >> KB);
        if constexpr (LY::custom) {
            int p = (u & ((1 << KB) - 1)) << B0;
            for (int i = 0; i < LOGN - KB; ++i)
                if ((q >> i) & 1) p |= 1 << LY::ord[K][i];
            return p;
        } else {
            return ((q >> B0) << (B0 + KB)) | ((u & ((1 << KB) - 1)) << B0) | (q & ((1 << B0) - 1));
        }
    }
    __host__ __device__ static constexpr int lane_of(int K, int p) {
        const int B0 = K * LOGE, KB = kb(K);
        if constexpr (LY::custom) {
            int q = 0;
            for (int i = 0; i < LOGN - KB; ++i)
                if ((p >> LY::ord[K][i]) & 1) q |= 1 << i;
            return q % G;
        } else {
            const int q = (p & ((1 << B0) - 1)) | ((p >> (B0 + KB)) << B0);
            return q % G;
        }
    }
    // the layout is a bijection of the row onto LDS entries below ROW, and phys / lane_of are inverse
    __host__ __device__ static constexpr bool layout_ok() {
        bool used[N + (N >> LOGE)] = {};
        for (int p = 0; p < N; ++p) {
            const int a = addr(p);
            if (a < 0 || a >= N + (N >> LOGE) || used[a]) return false;
            used[a] = true;
        }
        for (int K = 0; K < NPASS; ++K) {
            bool hit[N] = {};
            for (int lane = 0; lane < G; ++lane)
                for (int u = 0; u < E; ++u) {
                    const int p = phys_c(K, lane, u);
                    if (p < 0 || p >= N || hit[p] || lane_of(K, p) != lane) return false;
                    hit[p] = true;
                }
        }
        return true;
    }
    // Does the transition from grouping KA to KB move values between wavefronts?  Lanes of one row are consecutive
    // threads, so rows of G <= 64 lanes never do; otherwise a value crosses when its writer and reader lanes lie in
    // different 64-lane wavefronts.
    __host__ __device__ static constexpr bool cross(int KA, int KB_) {
        if (G <= 64) return false;
        for (int lane = 0; lane < G; ++lane)
            for (int u = 0; u < E; ++u)
                if (lane_of(KA, phys_c(KB_, lane, u)) / 64 != lane / 64) return true;
        return false;
    }
    // pass edges K <-> K+1 that cross wavefronts: always the top ones (B0 grows with K)
    __host__ __device__ static constexpr int first_cross() {
        for (int K = 0; K + 1 < NPASS; ++K)
            if (cross(K, K + 1) || cross(K + 1, K)) return K;
        return NPASS - 1;
    }
    static constexpr int FIRSTX = first_cross();
    __host__ __device__ static constexpr bool cross_on_top() {
        for (int K = 0; K + 1 < NPASS; ++K)
            if ((cross(K, K + 1) || cross(K + 1, K)) != (K >= FIRSTX) || cross(K, K + 1) != cross(K + 1, K))
                return false;
        return true;
    }
    static constexpr int NBUF = FIRSTX < NPASS - 1 ? 2 : 1;  // exchange buffers per row
    // LDS entry of the lane's element 0 in grouping K; element u sits addr(phys_c(K, 0, u)) further on (the lane's
    // bits and the element's bits are disjoint and addr is linear)
    template <int K>
    __device__ static __forceinline__ int base(int lane) {
        return addr(phys_c(K, lane, 0));
    }
    template <int K>
    __host__ __device__ static constexpr int off(int u) {
        return addr(phys_c(K, 0, u));
    }
};

// ---- one pass of levels LO..HI (compile-time) forward / adjoint, in registers.  Slot (gs, i, jl) of pass K holds
// the twiddle pair of the lane's pairs that share a group gs and low bits jl on the pass's level i.
template <typename T, int LOGN, int LOGE, int K, int LO, int HI>
__device__ __forceinline__ void pass_fwd(cplx<T> (&v)[1 << LOGE], const cplx<T>* tw, const cplx<T>* tz) {
    using C = Cfg<T, LOGN, LOGE>;
    constexpr int KB = C::kb(K);
    constexpr int B0 = K * LOGE;
    constexpr int PER = (1 << KB) - 1;
#pragma unroll
    for (int i = 0; i < KB; ++i) {
        const int m = B0 + i + 1;
        if (m < LO || m > HI) continue;
#pragma unroll
        for (int gs = 0; gs < (C::E >> KB); ++gs) {
#pragma unroll
            for (int jl = 0; jl < (1 << i); ++jl) {
                const bool one = B0 == 0 && jl == 0;  // t = 0: z^t = 1
                const int slot = gs * PER + (1 << i) - 1 + jl;
#pragma unroll
                for (int h = 0; h < (1 << (KB - 1 - i)); ++h) {
                    const int u0 = (gs << KB) | (h << (i + 1)) | jl;
                    const int u1 = u0 | (1 << i);
                    const cplx<T> f0 = v[u0], f1 = v[u1];
                    v[u0] = one ? add(f0, f1) : cfma(tw[slot], f1, f0);
                    v[u1] = cfma(tz[slot], f1, f0);
                }
            }
        }
    }
}

template <typename T, int LOGN, int LOGE, int K, int LO, int HI>
__device__ __forceinline__ void pass_adj(cplx<T> (&v)[1 << LOGE], const cplx<T>* tw, const cplx<T>* tz) {
    using C = Cfg<T, LOGN, LOGE>;
    constexpr int KB = C::kb(K);
    constexpr int B0 = K * LOGE;
    constexpr int PER = (1 << KB) - 1;
#pragma unroll
    for (int i = KB - 1; i >= 0; --i) {
        const int m = B0 + i + 1;
        if (m < LO || m > HI) continue;
#pragma unroll
        for (int gs = 0; gs < (C::E >> KB); ++gs) {
#pragma unroll
            for (int jl = 0; jl < (1 << i); ++jl) {
                const bool one = B0 == 0 && jl == 0;
                const int slot = gs * PER + (1 << i) - 1 + jl;
#pragma unroll
                for (int h = 0; h < (1 << (KB - 1 - i)); ++h) {
                    const int u0 = (gs << KB) | (h << (i + 1)) | jl;
                    const int u1 = u0 | (1 << i);
                    const cplx<T> g0 = v[u0], g1 = v[u1];
                    v[u0] = add(g0, g1);
                    v[u1] = one ? cfmac(tz[slot], g1, g0) : cfmac(tz[slot], g1, cmulc(tw[slot], g0));
                }
            }
        }
    }
}

// Orders a wavefront's LDS accesses before and after it for the compiler; the hardware already performs one
// wavefront's LDS operations in order, so no wait or barrier instruction is needed between its lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Regroup the row from grouping KA to KB through LDS.  An exchange that stays inside each wavefront needs only
// wave_sync: every lane writes and reads positions of its own wavefront's lanes, which no other wavefront touches in
// such an exchange.  An exchange across wavefronts writes, synchronises the block and reads; the edges K <-> K+1
// that cross (the top ones) alternate between two buffers (up: edge e uses buffer e % 2, down: (e + 1) % 2), so
// consecutive cross exchanges never share a buffer and the barrier of the next one orders every read of the previous
// one before the buffer is written again: one barrier per cross exchange instead of two.  X2 is the row's second
// buffer (null when the configuration has no cross exchange).
template <typename T, int LOGN, int LOGE, int KA, int KB_>
__device__ __forceinline__ void transition(cplx<T> (&v)[1 << LOGE], cplx<T>* X, cplx<T>* X2, int lane) {
    using C = Cfg<T, LOGN, LOGE>;
    constexpr bool X_WAVES = C::cross(KA, KB_);
    cplx<T>* B = X;
    if constexpr (X_WAVES) {
        constexpr int e = (KA < KB_ ? KA : KB_) - C::FIRSTX;  // cross edge index from the bottom
        constexpr int buf = KA < KB_ ? (e % 2) : ((e + 1) % 2);
        if constexpr (buf) B = X2;
    }
    cplx<T>* W = B + C::template base<KA>(lane);
    const cplx<T>* R = B + C::template base<KB_>(lane);
#pragma unroll
    for (int u = 0; u < C::E; ++u) W[C::template off<KA>(u)] = v[u];
    if constexpr (X_WAVES)
        __syncthreads();
    else
        wave_sync();
#pragma unroll
    for (int u = 0; u < C::E; ++u) v[u] = R[C::template off<KB_>(u)];
    if constexpr (!X_WAVES) wave_sync();
}

// Slots of the innermost level n (level KB - 1 of the last pass) and their pairs (u0, u0 | HS).
template <typename T, int LOGN, int LOGE>
struct Inner {
    static constexpr int K = Cfg<T, LOGN, LOGE>::NPASS - 1;
    static constexpr int KB = Cfg<T, LOGN, LOGE>::kb(K);
    static constexpr int PER = (1 << KB) - 1;
    static constexpr int HS = 1 << (KB - 1);
    static constexpr int GROUPS = (1 << LOGE) >> KB;
    __host__ __device__ static constexpr int slot(int gs, int jl) { return gs * PER + HS - 1 + jl; }
    __host__ __device__ static constexpr int u0(int gs, int jl) { return (gs << KB) | jl; }
};

// Level n forward and then adjoint on its pairs as one step, halved.  Forward maps (f0, f1) to g = (f0 + a f1,
// f0 + b f1) with a = z^t, b = z^(t+1); the adjoint maps g to (g0 + g1, conj(a) g0 + conj(b) g1) = (2 f0 + (a + b) f1,
// conj(a + b) f0 + (|a|^2 + |b|^2) f1), and |z| = 1 makes that 2 (f0 + h f1, conj(h) f0 + f1) with h = (a + b) / 2:
// 8 FMAs per pair instead of 18.  The result is half the Gram's.
template <typename T, int LOGN, int LOGE>
__device__ __forceinline__ void inner_half(cplx<T> (&v)[1 << LOGE], const cplx<T>* h) {
    using I = Inner<T, LOGN, LOGE>;
#pragma unroll
    for (int gs = 0; gs < I::GROUPS; ++gs) {
#pragma unroll
        for (int jl = 0; jl < I::HS; ++jl) {
            const int u0 = I::u0(gs, jl), u1 = u0 | I::HS;
            const cplx<T> f0 = v[u0], f1 = v[u1], hv = h[I::slot(gs, jl)];
            v[u0] = cfma(hv, f1, f0);
            v[u1] = cfmac(hv, f0, f1);
        }
    }
}

// levels LO..n forward then adjoint (the Gram B^H B), grouping 0 -> grouping 0, HALF the result: the innermost level
// runs as inner_half.  hh holds h = (z^t + z^(t+1)) / 2 at the innermost level's slots.
template <typename T, int LOGN, int LOGE, int K, int LO>
__device__ __forceinline__ void gram_half(cplx<T> (&v)[1 << LOGE], cplx<T>* X, cplx<T>* X2,
                                          const cplx<T> (*tw)[1 << LOGE], const cplx<T> (*tz)[1 << LOGE],
                                          const cplx<T>* hh, int lane) {
    using C = Cfg<T, LOGN, LOGE>;
    if constexpr (K + 1 < C::NPASS) {
        pass_fwd<T, LOGN, LOGE, K, LO, LOGN>(v, tw[K], tz[K]);
        transition<T, LOGN, LOGE, K, K + 1>(v, X, X2, lane);
        gram_half<T, LOGN, LOGE, K + 1, LO>(v, X, X2, tw, tz, hh, lane);
        transition<T, LOGN, LOGE, K + 1, K>(v, X, X2, lane);
        pass_adj<T, LOGN, LOGE, K, LO, LOGN>(v, tw[K], tz[K]);
    } else {
        pass_fwd<T, LOGN, LOGE, K, LO, LOGN - 1>(v, tw[K], tz[K]);
        inner_half<T, LOGN, LOGE>(v, hh);
        pass_adj<T, LOGN, LOGE, K, LO, LOGN - 1>(v, tw[K], tz[K]);
    }
}

// adjoint passes K..0 of levels LO..n starting from grouping K
template <typename T, int LOGN, int LOGE, int K, int LO>
__device__ __forceinline__ void adj_down(cplx<T> (&v)[1 << LOGE], cplx<T>* X, cplx<T>* X2,
                                         const cplx<T> (*tw)[1 << LOGE], const cplx<T> (*tz)[1 << LOGE], int lane) {
    pass_adj<T, LOGN, LOGE, K, LO, LOGN>(v, tw[K], tz[K]);
    if constexpr (K > 0) {
        transition<T, LOGN, LOGE, K, K - 1>(v, X, X2, lane);
        adj_down<T, LOGN, LOGE, K - 1, LO>(v, X, X2, tw, tz, lane);
    }
}

// forward passes K..NPASS-1 of levels LO..n starting from grouping K, ending in grouping NPASS-1
template <typename T, int LOGN, int LOGE, int K, int LO>
__device__ __forceinline__ void fwd_up(cplx<T> (&v)[1 << LOGE], cplx<T>* X, cplx<T>* X2,
                                       const cplx<T> (*tw)[1 << LOGE], const cplx<T> (*tz)[1 << LOGE], int lane) {
    pass_fwd<T, LOGN, LOGE, K, LO, LOGN>(v, tw[K], tz[K]);
    if constexpr (K + 1 < Cfg<T, LOGN, LOGE>::NPASS) {
        transition<T, LOGN, LOGE, K, K + 1>(v, X, X2, lane);
        fwd_up<T, LOGN, LOGE, K + 1, LO>(v, X, X2, tw, tz, lane);
    }
}

template <typename T, int LOGN, int LOGE, bool COMP, int LO, bool X64>
__global__ __launch_bounds__((Cfg<T, LOGN, LOGE>::NT), (Cfg<T, LOGN, LOGE>::MINW)) void rows_reg_kernel(RegArgs RA) {
    using C = Cfg<T, LOGN, LOGE>;
    static_assert(C::cross_on_top(), "the buffer alternation of transition() needs the cross-wavefront edges on top");
    static_assert(C::layout_ok(), "the exchange layout must be a bijection into the row's LDS buffer");
    constexpr int E = C::E;
    constexpr int NP = C::NPASS;
    constexpr int LAST = NP - 1;
    // global row type: the compute type, or complex128 with the explicit level in double (X64)
    using TR = std::conditional_t<X64, double, T>;
    static_assert(!X64 || LO == 2, "complex128 rows with a float core: the hybrid's op lists (explicit level 1)");
    const RowsArgs& A = RA.r;
    __shared__ cplx<T> Xs[C::NBUF * C::RPB * C::ROW];
    __shared__ int sdeg[C::RPB];

    // ---- dense workers: the leading blocks take the pass's dense tiles on the matrix cores while the row blocks
    // keep the vector ALUs busy (no barrier on this path)
    int bid = blockIdx.x;
    if (A.dn.workers > 0) {
        if (bid < A.dn.dblk) {
            const int w = xcd_block(bid, A.dn.dblk) * (C::NT / 64) + (int)(threadIdx.x >> 6);
            if (w < A.dn.workers) dense::worker<TR, 4, C::DENSE_BR, C::DENSE_BC>(LOGN, A.dn, w, threadIdx.x & 63, A.Y, A.Z);
            return;
        }
        bid -= A.dn.dblk;
    }

    const int tid = threadIdx.x;
    const int rib = tid / C::G;   // row in block
    const int lane = tid % C::G;  // lane within the row
    cplx<T>* X = Xs + rib * C::ROW;
    cplx<T>* X2 = C::NBUF > 1 ? X + C::RPB * C::ROW : nullptr;  // second exchange buffer (cross-wavefront edges)

    // ---- which rows.  With the frequency-grouped mapping a block runs the row groups sg0 .. sg1 - 1 (RA.iter groups
    // of RPB rows) of one frequency k one after another, so the per-frequency set-up below (the z^t table and the
    // twiddle pairs in registers) is done once for all of them; otherwise a block has one row per row slot.
    int k, sg0 = 0, sg1 = 1;
    long long row1 = 0;
    bool valid1 = true;
    if (RA.group_by_k) {
        const int nspb = (RA.SP + C::RPB - 1) / C::RPB;
        const int nbk = (nspb + RA.iter - 1) / RA.iter;  // blocks per frequency
        const int kk = bid / nbk;
        sg0 = (bid % nbk) * RA.iter;
        sg1 = min(sg0 + RA.iter, nspb);
        k = RA.korder ? RA.korder[kk] : kk;
    } else {
        row1 = (long long)bid * C::RPB + rib;
        valid1 = row1 < A.rows;
        if (!valid1) row1 = 0;
        k = (int)(row1 % A.period);
    }
    const c128 zk = A.z[k];

    // ---- table of z^t, t = 0 .. N/2, in the row's LDS buffer: copied from the plan's per-frequency table when there
    // is one, else built here (products of z^(2^b) in double, rounded to T, ztab_entry)
    if (A.ztab) {
        const cplx<T>* zt = reinterpret_cast<const cplx<T>*>(A.ztab) + (size_t)k * (C::N / 2 + 1);
        for (int t = lane; t <= C::N / 2; t += C::G) X[t] = zt[t];
    } else {
        c128 p2[LOGN];
        ztab_powers(zk, LOGN, p2);
        for (int t = lane; t <= C::N / 2; t += C::G) X[t] = cast<T>(ztab_entry(p2, LOGN, t));
    }
    __syncthreads();
    // ---- the lane's twiddle pairs (z^t, z^(t+1)) of every pass
    cplx<T> tw[NP][E], tz[NP][E];
#pragma unroll
    for (int K = 0; K < NP; ++K) {
        const int B0 = K * LOGE;
        const int KB = (LOGN - B0) < LOGE ? (LOGN - B0) : LOGE;
        const int PER = (1 << KB) - 1;
#pragma unroll
        for (int gs = 0; gs < (E >> KB); ++gs) {
            // the low B0 physical bits of the lane's pairs in this group (below every level of the pass)
            const int rq = brev_low(C::phys_c(K, lane, gs << KB) & ((1 << B0) - 1), B0);
#pragma unroll
            for (int i = 0; i < KB; ++i) {
#pragma unroll
                for (int jl = 0; jl < (1 << i); ++jl) {
                    const int t = (rq << i) | brev_low(jl, i);  // t < 2^(B0 + i) <= N/2
                    tw[K][gs * PER + (1 << i) - 1 + jl] = X[t];
                    tz[K][gs * PER + (1 << i) - 1 + jl] = X[t + 1];
                }
            }
        }
    }
    // h = (z^t + z^(t+1)) / 2 of the innermost level's slots, rounded once; the Gram uses h there instead of the
    // pair, which is rebuilt as (a, 2 h - a) for the plain forward / adjoint ops
    using IN = Inner<T, LOGN, LOGE>;
    cplx<T> hh[E];
#pragma unroll
    for (int gs = 0; gs < IN::GROUPS; ++gs)
#pragma unroll
        for (int jl = 0; jl < IN::HS; ++jl) {
            const int sl = IN::slot(gs, jl);
            hh[sl] = cast<T>(scale(0.5, add(cast<double>(tw[LAST][sl]), cast<double>(tz[LAST][sl]))));
        }
    // the twiddle pairs with the innermost level's z^(t+1) rebuilt from h (the original is not kept in registers)
    auto pairs_full = [&](cplx<T> (&tzf)[NP][E]) {
#pragma unroll
        for (int K = 0; K < NP; ++K)
#pragma unroll
            for (int u = 0; u < E; ++u) tzf[K][u] = tz[K][u];
#pragma unroll
        for (int gs = 0; gs < IN::GROUPS; ++gs)
#pragma unroll
            for (int jl = 0; jl < IN::HS; ++jl) {
                const int sl = IN::slot(gs, jl);
                tzf[LAST][sl] = sub(scale((T)2, hh[sl]), tw[LAST][sl]);
            }
    };
    __syncthreads();

    // natural (= grouping 0 physical) index of the lane's element 0; element u is x0 + u
    const int x0 = C::phys_c(0, lane, 0);
    static_assert(C::phys_c(0, 0, E - 1) == E - 1, "grouping 0 holds E consecutive values per lane");
    for (int sg = sg0; sg < sg1; ++sg) {  // block-uniform trip count: the ops' barriers stay uniform
        // the previous row's last exchange may still be read by other wavefronts, and a leading BT op writes the whole
        // buffer: a barrier between rows
        if (sg != sg0) __syncthreads();
        long long row = row1;
        bool valid = valid1;
        if (RA.group_by_k) {
            const int sp = sg * C::RPB + rib;
            valid = sp < RA.SP;
            row = (long long)(valid ? sp : 0) * A.period + k;
        }
        // ---- load the row in natural order (grouping 0: lane holds E consecutive values)
        cplx<T> v[E];
        const cplx<TR>* Yr = reinterpret_cast<const cplx<TR>*>(A.Y) + row * A.in_stride;
        int o0 = 0, o1 = RA.nmops;
        if constexpr (X64) {
            // level 1 pairs natural neighbours (2l, 2l + 1), which sit in one lane here, with twiddles (1, z):
            // (f0 + f1, f0 + z f1) in double from the complex128 row, rounded once to T (fastops.py:330)
            const bool l1 = RA.mop[0] == MOP_L1F;
#pragma unroll
            for (int u = 0; u < E; u += 2) {
                const c128 f0 = valid ? Yr[x0 + u] : mk(0.0, 0.0), f1 = valid ? Yr[x0 + u + 1] : mk(0.0, 0.0);
                v[u] = cast<T>(l1 ? add(f0, f1) : f0);
                v[u + 1] = cast<T>(l1 ? add(f0, mul(zk, f1)) : f1);
            }
            o0 = l1 ? 1 : 0;
            if (o1 > o0 && RA.mop[o1 - 1] == MOP_L1A) --o1;
        } else {
#pragma unroll
            for (int u = 0; u < E; ++u) v[u] = valid ? Yr[x0 + u] : mk((T)0, (T)0);
        }

        for (int o = o0; o < o1; ++o) {
            const int op = RA.mop[o];
            if (op == MOP_L1F) {
                pass_fwd<T, LOGN, LOGE, 0, 1, 1>(v, tw[0], tz[0]);
            } else if (op == MOP_L1A) {
                pass_adj<T, LOGN, LOGE, 0, 1, 1>(v, tw[0], tz[0]);
            } else if (op == MOP_GRAM) {
                gram_half<T, LOGN, LOGE, 0, LO>(v, X, X2, tw, tz, hh, lane);
#pragma unroll
                for (int u = 0; u < E; ++u) v[u] = scale((T)2, v[u]);  // exact
            } else if (op == MOP_BTNAT) {
                // natural -> physical order of levels 1..n through LDS, then the adjoint passes down to grouping 0
#pragma unroll
                for (int u = 0; u < E; ++u) X[C::addr(lam(x0 + u, LOGN))] = v[u];
                __syncthreads();
#pragma unroll
                for (int u = 0; u < E; ++u) v[u] = X[C::template base<LAST>(lane) + C::template off<LAST>(u)];
                __syncthreads();
                cplx<T> tzf[NP][E];
                pairs_full(tzf);
                adj_down<T, LOGN, LOGE, LAST, LO>(v, X, X2, tw, tzf, lane);
            } else if (op == MOP_BEND) {
                cplx<T> tzf[NP][E];
                pairs_full(tzf);
                fwd_up<T, LOGN, LOGE, 0, LO>(v, X, X2, tw, tzf, lane);
#pragma unroll
                for (int u = 0; u < E; ++u) X[C::template base<LAST>(lane) + C::template off<LAST>(u)] = v[u];
                __syncthreads();
#pragma unroll
                for (int u = 0; u < E; ++u) v[u] = X[C::addr(lam(x0 + u, LOGN))];
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
                // The iterate is zi; with COMP it is the unevaluated sum zi + zl, and the recurrence carries its
                // increment dl = z_j - z_(j-1) instead of the previous iterate: z_(j+1) = z_j + dl_j with
                // dl_j = (w a) r_j + (w - 1) dl_(j-1) (the reference's z + w (a r) + (w - 1)(z - z_prev),
                // cpp/oblique.cpp:93-97), each increment added to zi by TwoSum and its rounding error to zl.  dl starts
                // at z_0 - 0 (the zero start).  The residual uses zi alone: tau zl is below the rounding error of the
                // FP32 product M z.
                constexpr int EL = COMP ? E : 1;
                constexpr int EP = COMP ? 1 : E;
                cplx<T> y[E], zi[E], zp[EP], zl[EL], dl[EL];
                const double ak = valid ? A.a[k] : 0.0;
                const T aT = (T)ak, tau = (T)A.tau;
                // the step weights come from a table (no double division in the loop)
                const c128* wc = A.wcoef + A.wcoff[k];
#pragma unroll
                for (int u = 0; u < E; ++u) {
                    y[u] = v[u];
                    zi[u] = scale(aT, y[u]);
                    if constexpr (COMP) {
                        zl[u] = mk(fma(aT, y[u].x, -zi[u].x), fma(aT, y[u].y, -zi[u].y));
                        dl[u] = zi[u];
                    } else {
                        zp[u] = mk((T)0, (T)0);
                    }
                }
                for (int it = 1; it < dmax; ++it) {
                    const c128 wq = wc[it < d ? it : 0];  // loaded before the Gram: its latency hides behind it
#pragma unroll
                    for (int u = 0; u < E; ++u) v[u] = zi[u];
                    gram_half<T, LOGN, LOGE, 0, LO>(v, X, X2, tw, tz, hh, lane);  // v = M z / 2
                    if (it < d) {
                        // z <- z + (w a) r + (w - 1)(z - z_prev), r = y - M z - tau z, y - M z as one rounding of
                        // y - 2 v
                        const T wa = (T)wq.x, w0 = (T)wq.y;
#pragma unroll
                        for (int u = 0; u < E; ++u) {
                            if constexpr (COMP) {
                                const cplx<T> r = sub(mk(fma((T)-2, v[u].x, y[u].x), fma((T)-2, v[u].y, y[u].y)),
                                                      scale(tau, zi[u]));
                                dl[u] = add(scale(wa, r), scale(w0, dl[u]));
                                cplx<T> hi, err;
                                two_sum(zi[u], dl[u], hi, err);
                                zi[u] = hi;
                                zl[u] = add(zl[u], err);
                            } else {
                                const cplx<T> r = sub(mk(fma((T)-2, v[u].x, y[u].x), fma((T)-2, v[u].y, y[u].y)),
                                                      scale(tau, zi[u]));
                                const cplx<T> zn = add(add(zi[u], scale(wa, r)), scale(w0, sub(zi[u], zp[u])));
                                zp[u] = zi[u];
                                zi[u] = zn;
                            }
                        }
                    }
                }
                if (anyd) {
                    // exact Tikhonov filter z = F y in double for dense rows (F^T stored: coalesced over the lane's E)
#pragma unroll
                    for (int u = 0; u < E; ++u) X[C::addr(x0 + u)] = y[u];
                    __syncthreads();
                    if (dense) {
                        const c128* F = A.dFT + (long long)slot * C::N * C::N;
                        c128 acc[E];
#pragma unroll
                        for (int u = 0; u < E; ++u) acc[u] = mk(0.0, 0.0);
                        for (int q = 0; q < C::N; ++q) {
                            const c128 yq = cast<double>(X[C::addr(q)]);
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
            cplx<TR>* Zr = reinterpret_cast<cplx<TR>*>(A.Z) + row * A.out_stride;
            if constexpr (X64) {
                // the adjoint of level 1, (g0 + g1, g0 + conj(z) g1), in double on the float results (fastops.py:330)
                const bool l1 = o1 < RA.nmops;
#pragma unroll
                for (int u = 0; u < E; u += 2) {
                    const c128 g0 = cast<double>(v[u]), g1 = cast<double>(v[u + 1]);
                    Zr[x0 + u] = l1 ? add(g0, g1) : g0;
                    Zr[x0 + u + 1] = l1 ? add(g0, mul(conj(zk), g1)) : g1;
                }
            } else {
#pragma unroll
                for (int u = 0; u < E; ++u) Zr[x0 + u] = v[u];
            }
        }
    }  // row groups
}

template <typename T, int LOGN, int LOGE, bool COMP, int LO, bool X64>
int launch_cfg(const RegArgs& RA, hipStream_t s) {
    using C = Cfg<T, LOGN, LOGE>;
    RegArgs RB = RA;
    long long nblk;
    if (RA.group_by_k) {
        // Float rows: several row groups of a frequency per block share its set-up, as long as the launch keeps at
        // least 8 wavefronts of rows per SIMD (8 x 1024 on MI355X) to fill the chip and balance its tail (N=1024,
        // 8 sinograms: 1166 -> 1112 us per launch; double rows measured slower with it, 125 -> 136 us)
        const int nspb = (RA.SP + C::RPB - 1) / C::RPB;
        RB.iter = 1;
        for (int it = sizeof(T) == 4 ? 4 : 1; it > 1; it /= 2)
            if ((long long)RA.nk * ((nspb + it - 1) / it) * (C::NT / 64) >= 8 * 1024) {
                RB.iter = it;
                break;
            }
        nblk = (long long)RA.nk * ((nspb + RB.iter - 1) / RB.iter);
    } else {
        RB.iter = 1;
        nblk = (RA.r.rows + C::RPB - 1) / C::RPB;
    }
    if (RA.r.dn.n_dense > 0) {
        // dense workers in leading blocks of C::NT / 64 waves, padded to a multiple of the XCD count
        const bool blk = dense::row_tiles(RA.r.dn) >= C::DENSE_BR && (1 << LOGN) / 16 >= C::DENSE_BC;
        RB.r.dn.rt = blk ? C::DENSE_BR : 1;
        RB.r.dn.ct = blk ? C::DENSE_BC : 1;
        if (dense::tiles(LOGN, RB.r.dn) > 0x7fff0000LL) {  // the workers step through units with int counters
            set_error("rows_reg: too many dense tiles");
            return E_INVALID;
        }
        RB.r.dn.workers = dense_workers(LOGN, sizeof(T) == 8 ? PREC_F64 : PREC_F32, RB.r.dn);
        const int wpb = C::NT / 64;
        RB.r.dn.dblk = (int)xcd_grid((RB.r.dn.workers + wpb - 1) / wpb);
        nblk += RB.r.dn.dblk;
    } else {
        RB.r.dn.workers = RB.r.dn.dblk = 0;
    }
    if (nblk < 1) return OK;
    if (nblk > 0x7fffffffLL) {
        set_error("rows_reg: too many blocks");
        return E_INVALID;
    }
    if (RA.r.ev_start)
        hipExtLaunchKernelGGL((rows_reg_kernel<T, LOGN, LOGE, COMP, LO, X64>), dim3((unsigned)nblk), dim3(C::NT), 0, s,
                              RA.r.ev_start, RA.r.ev_stop, 0, RB);
    else
        hipLaunchKernelGGL((rows_reg_kernel<T, LOGN, LOGE, COMP, LO, X64>), dim3((unsigned)nblk), dim3(C::NT), 0, s, RB);
    return launch_check("rows_reg_kernel");
}

template <typename T, int LOGN, bool COMP, int LO, bool X64>
int launch_n(const RegArgs& RA, int loge, hipStream_t s) {
    if constexpr (LOGN >= 2) {
        if (loge == 2) return launch_cfg<T, LOGN, 2, COMP, LO, X64>(RA, s);
    }
    if constexpr (LOGN >= 3) {
        if (loge == 3) return launch_cfg<T, LOGN, 3, COMP, LO, X64>(RA, s);
    }
    if constexpr (LOGN >= 4) {
        if (loge == 4) return launch_cfg<T, LOGN, 4, COMP, LO, X64>(RA, s);
    }
    set_error("rows_reg: unsupported elements-per-lane");
    return E_INVALID;
}

template <typename T, bool COMP, int LO, bool X64 = false>
int launch_t(const RegArgs& RA, int loge, hipStream_t s) {
    switch (RA.r.n) {
        case 5: return launch_n<T, 5, COMP, LO, X64>(RA, loge, s);
        case 6: return launch_n<T, 6, COMP, LO, X64>(RA, loge, s);
        case 7: return launch_n<T, 7, COMP, LO, X64>(RA, loge, s);
        case 8: return launch_n<T, 8, COMP, LO, X64>(RA, loge, s);
        case 9: return launch_n<T, 9, COMP, LO, X64>(RA, loge, s);
        case 10: return launch_n<T, 10, COMP, LO, X64>(RA, loge, s);
        case 11: return launch_n<T, 11, COMP, LO, X64>(RA, loge, s);
        default: set_error("rows_reg: n must be 5..11"); return E_INVALID;
    }
}

}  // namespace rreg

}  // namespace adrt
