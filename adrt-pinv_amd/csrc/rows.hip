// Fused row kernel (see rows.hpp for the in-place index scheme) and its host launcher.
#include <cstdlib>

#include "rows.hpp"

namespace adrt {

namespace {

__device__ __forceinline__ int rotl_low(int x, int k) {
    if (k <= 1) return x;
    const int mask = (1 << k) - 1;
    const int low = x & mask;
    return (x & ~mask) | (((low << 1) | (low >> (k - 1))) & mask);
}

// Logical index held by physical slot p after forward levels lo..hi (identity when hi < lo).
__device__ __forceinline__ int lam_perm(int p, int lo, int hi) {
    if (hi < lo) return p;
    if (lo <= 2) {
        if (hi <= 1) return p;
        const int mask = (1 << hi) - 1;
        return (p & ~mask) | (int)(__brev((unsigned)(p & mask)) >> (32 - hi));
    }
    for (int k = lo; k <= hi; ++k) p = rotl_low(p, k);
    return p;
}

template <typename T>
__device__ __forceinline__ void level_fwd(cplx<T>* X, const cplx<T>* tw, cplx<T> z, int N, int m, int lo, int tid,
                                          int nt) {
    const int w = 1 << (m - 1);
    for (int q = tid; q < (N >> 1); q += nt) {
        const int p0 = ((q >> (m - 1)) << m) | (q & (w - 1));
        const int p1 = p0 + w;
        const int t = lam_perm(p0, lo, m - 1) & (w - 1);
        const cplx<T> f0 = X[p0], f1 = X[p1];
        const cplx<T> u = mul(tw[t], f1);
        X[p0] = add(f0, u);
        X[p1] = add(f0, mul(u, z));
    }
}

template <typename T>
__device__ __forceinline__ void level_adj(cplx<T>* X, const cplx<T>* tw, cplx<T> zc, int N, int m, int lo, int tid,
                                          int nt) {
    const int w = 1 << (m - 1);
    for (int q = tid; q < (N >> 1); q += nt) {
        const int p0 = ((q >> (m - 1)) << m) | (q & (w - 1));
        const int p1 = p0 + w;
        const int t = lam_perm(p0, lo, m - 1) & (w - 1);
        const cplx<T> g0 = X[p0], g1 = X[p1];
        X[p0] = add(g0, g1);
        X[p1] = cmulc(tw[t], add(g0, mul(zc, g1)));
    }
}

template <typename T>
__device__ __forceinline__ void levels_fwd(cplx<T>* X, const cplx<T>* tw, cplx<T> z, int N, int lo, int hi, int tid,
                                           int nt) {
    for (int m = lo; m <= hi; ++m) {
        level_fwd(X, tw, z, N, m, lo, tid, nt);
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ void levels_adj(cplx<T>* X, const cplx<T>* tw, cplx<T> zc, int N, int lo, int hi, int tid,
                                           int nt) {
    for (int m = hi; m >= lo; --m) {
        level_adj(X, tw, zc, N, m, lo, tid, nt);
        __syncthreads();
    }
}

// Permute the LDS row between natural order and the physical order of levels lo..hi.  to_phys: X[p] <- X[Lambda(p)].
template <typename T, int NT, int E>
__device__ __forceinline__ void relayout(cplx<T>* X, int N, int lo, int hi, bool to_phys, int tid) {
    cplx<T> v[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int p = tid + NT * j;
        if (p < N) v[j] = X[to_phys ? lam_perm(p, lo, hi) : p];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int p = tid + NT * j;
        if (p < N) X[to_phys ? p : lam_perm(p, lo, hi)] = v[j];
    }
    __syncthreads();
}

// One row per block.  Tg: global row type, T: compute type.  NT threads, E = ceil(N / NT) natural elements per thread.
template <typename Tg, typename T, int NT, int E>
__global__ __launch_bounds__(NT) void rows_kernel(RowsArgs A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = A.N;
    c128* pw2 = reinterpret_cast<c128*>(smem_raw);                // z^(2^j), j < n (double)
    cplx<T>* X = reinterpret_cast<cplx<T>*>(smem_raw + 16 * sizeof(c128));
    cplx<T>* tw = X + N;                                           // N/2 twiddles z^t
    const int tid = threadIdx.x;

    for (long long row = blockIdx.x; row < A.rows; row += gridDim.x) {
        const int k = (int)(row % A.period);
        if (A.skip_dense && A.dslot && A.dslot[k] >= 0) continue;  // block-uniform: one row per block
        const c128 zk = A.z[k];
        // twiddles z^t = prod over set bits j of t of z^(2^j), in double
        if (tid == 0) {
            c128 p = zk;
            for (int j = 0; j < A.n; ++j) {
                pw2[j] = p;
                p = mul(p, p);
            }
        }
        __syncthreads();
        for (int t = tid; t < (N >> 1); t += NT) {
            c128 v = mk(1.0, 0.0);
            for (int j = 0; (t >> j) != 0; ++j)
                if ((t >> j) & 1) v = mul(v, pw2[j]);
            tw[t] = cast<T>(v);
        }
        const cplx<T> z = cast<T>(zk);
        const cplx<T> zc = conj(z);
        // load the row (natural order)
        const cplx<Tg>* Yr = reinterpret_cast<const cplx<Tg>*>(A.Y) + row * A.in_stride;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int p = tid + NT * j;
            if (p < N) X[p] = cast<T>(Yr[p]);
        }
        __syncthreads();

        int lay_lo = 0, lay_hi = 0;  // current physical order (0, 0) = natural
        for (int o = 0; o < A.nops; ++o) {
            const RowOp op = A.ops[o];
            if (op.kind == OP_B) {
                if (lay_hi != 0) {
                    relayout<T, NT, E>(X, N, lay_lo, lay_hi, false, tid);
                    lay_lo = lay_hi = 0;
                }
                levels_fwd(X, tw, z, N, op.lo, op.hi, tid, NT);
                lay_lo = op.lo;
                lay_hi = op.hi;
            } else if (op.kind == OP_BT) {
                if (!(lay_lo == op.lo && lay_hi == op.hi)) {
                    if (lay_hi != 0) relayout<T, NT, E>(X, N, lay_lo, lay_hi, false, tid);
                    relayout<T, NT, E>(X, N, op.lo, op.hi, true, tid);
                }
                levels_adj(X, tw, zc, N, op.lo, op.hi, tid, NT);
                lay_lo = lay_hi = 0;
            } else {  // OP_P: natural order in and out
                if (lay_hi != 0) {
                    relayout<T, NT, E>(X, N, lay_lo, lay_hi, false, tid);
                    lay_lo = lay_hi = 0;
                }
                cplx<T> y[E], zi[E], zp[E];
#pragma unroll
                for (int j = 0; j < E; ++j) {
                    const int p = tid + NT * j;
                    if (p < N) y[j] = X[p];
                }
                const int slot = A.dslot ? A.dslot[k] : -1;
                if (slot >= 0) {
                    // exact Tikhonov filter z = F y in double; F^T is stored so thread i streams column i
                    const c128* F = A.dFT + (long long)slot * N * N;
                    c128 acc[E];
#pragma unroll
                    for (int j = 0; j < E; ++j) acc[j] = mk(0.0, 0.0);
                    for (int q = 0; q < N; ++q) {
                        const c128 yq = cast<double>(X[q]);
#pragma unroll
                        for (int j = 0; j < E; ++j) {
                            const int p = tid + NT * j;
                            if (p < N) acc[j] = add(acc[j], mul(F[(long long)q * N + p], yq));
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < E; ++j) {
                        const int p = tid + NT * j;
                        if (p < N) X[p] = cast<T>(acc[j]);
                    }
                    __syncthreads();
                } else {
                    const double ak = A.a[k], ck = A.c[k];
                    const int d = A.deg[k];
                    const T tau = (T)A.tau;
                    // compensated FP32: the iterate is zi + zl, increments added by TwoSum
                    const bool comp = A.comp && sizeof(T) == sizeof(float);
                    cplx<T> zl[E], zpl[E];
#pragma unroll
                    for (int j = 0; j < E; ++j) {
                        zi[j] = scale((T)ak, y[j]);
                        zp[j] = mk((T)0, (T)0);
                        zl[j] = comp ? mk(fma((T)ak, y[j].x, -zi[j].x), fma((T)ak, y[j].y, -zi[j].y)) : mk((T)0, (T)0);
                        zpl[j] = mk((T)0, (T)0);
                    }
                    double om = 1.0;
                    for (int it = 1; it < d; ++it) {
                        const double omj = (it == 1) ? 1.0 / (1.0 - ck * ck / 2.0) : 1.0 / (1.0 - ck * ck * om / 4.0);
                        __syncthreads();
#pragma unroll
                        for (int j = 0; j < E; ++j) {
                            const int p = tid + NT * j;
                            if (p < N) X[p] = zi[j];
                        }
                        __syncthreads();
                        levels_fwd(X, tw, z, N, op.lo, op.hi, tid, NT);
                        levels_adj(X, tw, zc, N, op.lo, op.hi, tid, NT);
                        const T w1 = (T)omj, w0 = (T)(omj - 1.0), aT = (T)ak;
#pragma unroll
                        for (int j = 0; j < E; ++j) {
                            const int p = tid + NT * j;
                            if (p < N && comp) {
                                const cplx<T> r = sub(sub(sub(y[j], X[p]), scale(tau, zi[j])), scale(tau, zl[j]));
                                const cplx<T> dz = add(sub(zi[j], zp[j]), sub(zl[j], zpl[j]));
                                cplx<T> hi, err;
                                two_sum(zi[j], add(scale(w1, scale(aT, r)), scale(w0, dz)), hi, err);
                                zp[j] = zi[j];
                                zpl[j] = zl[j];
                                zi[j] = hi;
                                zl[j] = add(zl[j], err);
                            } else if (p < N) {
                                const cplx<T> r = sub(sub(y[j], X[p]), scale(tau, zi[j]));
                                const cplx<T> zn = add(add(zi[j], scale(w1, scale(aT, r))), scale(w0, sub(zi[j], zp[j])));
                                zp[j] = zi[j];
                                zi[j] = zn;
                            }
                        }
                        om = omj;
                    }
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < E; ++j) {
                        const int p = tid + NT * j;
                        if (p < N) X[p] = comp ? add(zi[j], zl[j]) : zi[j];
                    }
                    __syncthreads();
                }
            }
        }
        if (lay_hi != 0) relayout<T, NT, E>(X, N, lay_lo, lay_hi, false, tid);
        cplx<Tg>* Zr = reinterpret_cast<cplx<Tg>*>(A.Z) + row * A.out_stride;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int p = tid + NT * j;
            if (p < N) Zr[p] = cast<Tg>(X[p]);
        }
        __syncthreads();
    }
}

template <typename Tg, typename T, int NT, int E>
int launch_one(const RowsArgs& A, hipStream_t s) {
    const size_t lds = sizeof(cplx<T>) * (A.N + (A.N >> 1) + 1) + sizeof(c128) * 16 + 16;
    long long grid = A.rows < (1LL << 20) ? A.rows : (1LL << 20);
    if (grid < 1) return OK;
    if (A.ev_start)
        hipExtLaunchKernelGGL((rows_kernel<Tg, T, NT, E>), dim3((unsigned)grid), dim3(NT), (std::uint32_t)lds, s,
                              A.ev_start, A.ev_stop, 0, A);
    else
        hipLaunchKernelGGL((rows_kernel<Tg, T, NT, E>), dim3((unsigned)grid), dim3(NT), lds, s, A);
    return launch_check("rows_kernel");
}

template <typename Tg, typename T>
int dispatch(const RowsArgs& A, hipStream_t s) {
    const int N = A.N;
    if (N <= 64) return launch_one<Tg, T, 64, 1>(A, s);
    if (N == 128) return launch_one<Tg, T, 64, 2>(A, s);
    if (N == 256) return launch_one<Tg, T, 64, 4>(A, s);
    if (N == 512) return launch_one<Tg, T, 128, 4>(A, s);
    if (N == 1024) return launch_one<Tg, T, 256, 4>(A, s);
    if (N == 2048) return launch_one<Tg, T, 256, 8>(A, s);
    set_error("rows: N must be a power of two <= 2048");
    return E_INVALID;
}

}  // namespace

// Per-(n, class) choice of the register kernel's elements per lane: 0 = built-in default, -1 = generic kernel,
// 2..4 = log2 elements per lane.  Classes: 0 FP64 passes, 1 FP32 passes, 2 the FP64 refinement passes of FP32 plans
// (their batch depth differs from a plain FP64 plan's; unset, they follow class 0).  Set by the autotuner
// (adrt_rows_config_set) or the environment variable ADRT_ROWS_LOGE (one value for every size, read once).
static int g_rows_cfg[3][12];
static int g_rows_env = -100;

int rows_config_get(int n, int prec) {
    if (g_rows_env == -100) {
        const char* e = getenv("ADRT_ROWS_LOGE");
        g_rows_env = e ? atoi(e) : 0;
    }
    if (n < 0 || n > 11 || prec < 0 || prec > 2) return 0;
    if (g_rows_cfg[prec][n] != 0) return g_rows_cfg[prec][n];
    if (prec == 2) return rows_config_get(n, 0);
    return g_rows_env;
}

int rows_config_set(int n, int prec, int loge) {
    ADRT_REQUIRE(n >= 0 && n <= 11 && prec >= 0 && prec <= 2, "rows_config_set: bad n or precision class");
    ADRT_REQUIRE(loge >= -1 && loge <= 4 && loge != 1, "rows_config_set: loge must be -1, 0, 2, 3 or 4");
    g_rows_cfg[prec][n] = loge;
    return OK;
}

// classes 0 / 1: FP64 / FP32 passes; 2 / 3: the same with one sinogram (two rows per frequency), where the W stream
// dominates and more dense waves can pay (0 there: as class 0 / 1)
static int g_dense_cfg[4][12];

int dense_config_get(int n, int prec) {
    if (n < 0 || n > 11 || prec < 0 || prec > 3) return 0;
    if (prec >= 2 && g_dense_cfg[prec][n] == 0) return g_dense_cfg[prec - 2][n];
    return g_dense_cfg[prec][n];
}

int dense_class(int prec, const DenseArgs& dn) { return prec + (dn.SP <= 2 ? 2 : 0); }

int dense_config_set(int n, int prec, int mode) {
    ADRT_REQUIRE(n >= 0 && n <= 11 && prec >= 0 && prec <= 3, "dense_config_set: bad n or class");
    ADRT_REQUIRE(mode >= -1 && mode <= 8, "dense_config_set: mode must be -1 .. 8");
    g_dense_cfg[prec][n] = mode;
    return OK;
}

int check_ops(int n, const RowOp* ops, int nops) {
    ADRT_REQUIRE(nops >= 0 && nops <= MAX_OPS, "op list longer than 8");
    for (int i = 0; i < nops; ++i) {
        const RowOp& o = ops[i];
        ADRT_REQUIRE(o.kind == OP_B || o.kind == OP_BT || o.kind == OP_P, "op kind must be B, BT or P");
        ADRT_REQUIRE(1 <= o.lo && o.lo <= o.hi && o.hi <= n, "op levels must satisfy 1 <= lo <= hi <= n");
    }
    return OK;
}

int launch_rows(const RowsArgs& A, int prec, hipStream_t s, int* dense_done) {
    if (dense_done) *dense_done = 0;
    ADRT_REQUIRE(A.n >= 1 && A.n <= 11 && A.N == (1 << A.n), "rows: need 1 <= n <= 11 and N = 2^n");
    ADRT_TRY(check_ops(A.n, A.ops, A.nops));
    ADRT_REQUIRE(A.period >= 1, "rows: period must be positive");
    ADRT_REQUIRE(A.dn.n_dense == 0 || (A.skip_dense && A.Z != A.Y && dense_done),
                 "rows: dense tiles need skip_dense, distinct input and output rows and a completion flag");
    const int cls = A.cfg_class >= 0 ? A.cfg_class : prec;
    if (A.io64) {
        // complex128 rows with a float core: the register kernel only (the generic kernel's rows are the compute type)
        ADRT_REQUIRE(A.loge >= 0, "rows: complex128 rows with a float core need the register kernel");
        const int cfg = rows_config_get(A.n, cls);
        const int loge = A.loge > 0 ? A.loge : (cfg > 0 ? cfg : default_loge(A.n, prec));
        ADRT_REQUIRE(loge > 0, "rows: complex128 rows with a float core need 32 <= N <= 2048");
        const bool fuse = A.dn.n_dense > 0 && dense_config_get(A.n, dense_class(prec, A.dn)) >= 0;
        RowsArgs R = A;
        if (!fuse) R.dn = DenseArgs{};
        const int st = launch_rows_reg(R, prec, loge, s);
        if (st == E_INVALID) set_error("rows: the register kernel does not take this op list or size with complex128 rows");
        if (st == OK && dense_done) *dense_done = fuse;
        return st;
    }
    if (A.loge >= 0 && rows_config_get(A.n, cls) >= 0) {
        const int cfg = rows_config_get(A.n, cls);
        const int loge = A.loge > 0 ? A.loge : (cfg != 0 ? cfg : default_loge(A.n, prec));
        if (loge > 0) {
            const bool fuse = A.dn.n_dense > 0 && dense_config_get(A.n, dense_class(prec, A.dn)) >= 0;
            RowsArgs R = A;
            if (!fuse) R.dn = DenseArgs{};
            const int st = launch_rows_reg(R, prec, loge, s);
            if (st == OK) {
                if (dense_done) *dense_done = fuse;
                return OK;
            }
            if (st != E_INVALID) return st;
        }
    }
    if (prec == PREC_F64) return dispatch<double, double>(A, s);
    if (prec == PREC_F32) return dispatch<float, float>(A, s);
    set_error("rows: unknown precision");
    return E_INVALID;
}

}  // namespace adrt
