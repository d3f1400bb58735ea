// Shared definitions of the ADRT HIP library: complex arithmetic, error reporting, launch checks.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace adrt {

// ---------------------------------------------------------------- complex values in registers / LDS / HBM
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};
using c64 = cplx<float>;
using c128 = cplx<double>;

template <typename T>
__host__ __device__ inline cplx<T> mk(T x, T y) { return cplx<T>{x, y}; }
template <typename T>
__host__ __device__ inline cplx<T> add(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ inline cplx<T> sub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ inline cplx<T> mul(cplx<T> a, cplx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
__host__ __device__ inline cplx<T> scale(T s, cplx<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__host__ __device__ inline cplx<T> conj(cplx<T> a) { return {a.x, -a.y}; }
// conj(a) * b
template <typename T>
__host__ __device__ inline cplx<T> cmulc(cplx<T> a, cplx<T> b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }
template <typename T, typename U>
__host__ __device__ inline cplx<T> cast(cplx<U> a) { return {(T)a.x, (T)a.y}; }
// c + a * b and c + conj(a) * b as two FMAs per component
template <typename T>
__device__ inline cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> c) {
    return {fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
template <typename T>
__device__ inline cplx<T> cfmac(cplx<T> a, cplx<T> b, cplx<T> c) {
    return {fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y))};
}

// Error-free sum (Knuth's TwoSum): s = fl(a + b) and e with a + b = s + e exactly in round-to-nearest.  Additions
// only, so contraction into FMAs cannot alter it.
template <typename T>
__host__ __device__ inline void two_sum(T a, T b, T& s, T& e) {
    s = a + b;
    const T bv = s - a;
    e = (a - (s - bv)) + (b - bv);
}
template <typename T>
__host__ __device__ inline void two_sum(cplx<T> a, cplx<T> b, cplx<T>& s, cplx<T>& e) {
    two_sum(a.x, b.x, s.x, e.x);
    two_sum(a.y, b.y, s.y, e.y);
}

// ---------------------------------------------------------------- XCD-grouped block numbering
// The dispatcher deals consecutive hardware block ids to the device's 8 XCDs in turn, and each XCD has its own L2.
// Launching xcd_grid(nblk) blocks and taking logical block xcd_block(blockIdx.x, nblk) gives every XCD a contiguous
// range of logical blocks, so blocks that read the same data (one image's tiles, one frequency's rows) share an L2.
// xcd_block returns -1 for the few padding blocks of the last range (they must return before any barrier use).
constexpr int XCDS = 8;
__device__ __forceinline__ int xcd_block(int hw, int nblk) {
    const int per = (nblk + XCDS - 1) / XCDS;
    const int lb = (hw % XCDS) * per + hw / XCDS;
    return lb < nblk ? lb : -1;
}
inline unsigned xcd_grid(long long nblk) { return (unsigned)(((nblk + XCDS - 1) / XCDS) * XCDS); }

// ---------------------------------------------------------------- error reporting (thread-local last error)
void set_error(const std::string& msg);
const char* last_error_cstr();

enum Status : int {
    OK = 0,
    E_INVALID = -1,  // bad argument (size, level range, null pointer, op list)
    E_HIP = -2,      // HIP runtime error
    E_NOMEM = -3,    // device memory budget exceeded or allocation failed
    E_STATE = -4,    // object used in the wrong state (e.g. launching an uncaptured graph)
};

#define ADRT_HIP(call)                                                                                      \
    do {                                                                                                    \
        hipError_t _e = (call);                                                                             \
        if (_e != hipSuccess) {                                                                             \
            ::adrt::set_error(std::string(#call " failed: ") + hipGetErrorString(_e) + " at " __FILE__ ":" + \
                              std::to_string(__LINE__));                                                    \
            return ::adrt::E_HIP;                                                                           \
        }                                                                                                   \
    } while (0)

#define ADRT_REQUIRE(cond, msg)                       \
    do {                                              \
        if (!(cond)) {                                \
            ::adrt::set_error(std::string(msg));      \
            return ::adrt::E_INVALID;                 \
        }                                             \
    } while (0)

#define ADRT_TRY(expr)              \
    do {                            \
        int _s = (expr);            \
        if (_s != ::adrt::OK) return _s; \
    } while (0)

inline int launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("launch of ") + what + " failed: " + hipGetErrorString(e));
        return E_HIP;
    }
    return OK;
}

// ---------------------------------------------------------------- row operations
enum OpKind : int { OP_B = 0, OP_BT = 1, OP_P = 2 };
constexpr int MAX_OPS = 8;
struct RowOp {
    int kind, lo, hi;
};

// Precision of the row engine.  F64: every row op in double.  F32: butterflies, Chebyshev steps and the
// offset FFTs in float; the dense table and everything in image space in double.  With compensation (RowsArgs::comp)
// the FP32 Chebyshev iterate is carried as an unevaluated sum hi + lo, every increment added by TwoSum.
enum Precision : int { PREC_F64 = 0, PREC_F32 = 1 };

}  // namespace adrt
