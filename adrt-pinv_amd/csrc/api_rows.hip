// C ABI, part 1: errors, device memory and streams, the reference-compatible row kernels on host buffers
// (bfly_rows / cheb_rows / cert_rowsum / omp_max_threads, cpp/oblique.cpp:53-139), and the device-resident
// row plan used by the row-engine binding.
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/adrt_hip.h"
#include "rows.hpp"

namespace adrt {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }

namespace {

// ---------------------------------------------------------------- scratch for the host-buffer entry points
// The reference-compatible calls copy host buffers in and out; their device buffers are cached per process
// and grown on demand (one call at a time: they serialise on this mutex).
struct HostCallScratch {
    std::mutex mu;
    void* buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t cap[6] = {0, 0, 0, 0, 0, 0};
    hipStream_t stream = nullptr;
};
HostCallScratch& scratch() {
    static HostCallScratch s;
    return s;
}

int ensure(int i, size_t bytes) {
    HostCallScratch& s = scratch();
    if (s.cap[i] >= bytes) return OK;
    if (s.buf[i]) ADRT_HIP(hipFree(s.buf[i]));
    s.buf[i] = nullptr;
    s.cap[i] = 0;
    size_t want = bytes < 256 ? 256 : bytes;
    ADRT_HIP(hipMalloc(&s.buf[i], want));
    s.cap[i] = want;
    return OK;
}

int ensure_stream() {
    HostCallScratch& s = scratch();
    if (!s.stream) ADRT_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    return OK;
}

void report(int st, const char* what) {
    if (st != OK) std::fprintf(stderr, "[adrt_hip] %s: %s\n", what, last_error_cstr());
}

int bfly_rows_impl(int n, int K, const void* z, void* X, int m_lo, int m_hi, int transpose) {
    ADRT_REQUIRE(n >= 1 && n <= 11, "bfly_rows: need 1 <= n <= 11");
    ADRT_REQUIRE(K >= 0, "bfly_rows: K must be >= 0");
    ADRT_REQUIRE(1 <= m_lo && m_lo <= m_hi && m_hi <= n, "bfly_rows: need 1 <= m_lo <= m_hi <= n");
    ADRT_REQUIRE(transpose == 0 || transpose == 1, "bfly_rows: transpose must be 0 or 1");
    if (K == 0) return OK;
    ADRT_REQUIRE(z && X, "bfly_rows: null pointer");
    const int N = 1 << n;
    HostCallScratch& s = scratch();
    std::lock_guard<std::mutex> g(s.mu);
    ADRT_TRY(ensure_stream());
    const size_t rb = sizeof(c128) * (size_t)K * N;
    ADRT_TRY(ensure(0, sizeof(c128) * (size_t)K));
    ADRT_TRY(ensure(1, rb));
    ADRT_HIP(hipMemcpyAsync(s.buf[0], z, sizeof(c128) * K, hipMemcpyHostToDevice, s.stream));
    ADRT_HIP(hipMemcpyAsync(s.buf[1], X, rb, hipMemcpyHostToDevice, s.stream));
    RowsArgs A{};
    A.Y = s.buf[1];
    A.Z = s.buf[1];
    A.rows = K;
    A.in_stride = A.out_stride = N;
    A.n = n;
    A.N = N;
    A.period = K;
    A.z = (const c128*)s.buf[0];
    A.nops = 1;
    A.ops[0] = RowOp{transpose ? OP_BT : OP_B, m_lo, m_hi};
    ADRT_TRY(launch_rows(A, PREC_F64, s.stream));
    ADRT_HIP(hipMemcpyAsync(X, s.buf[1], rb, hipMemcpyDeviceToHost, s.stream));
    ADRT_HIP(hipStreamSynchronize(s.stream));
    return OK;
}

int cheb_rows_impl(int n, int K, const void* z, const int* deg, const double* a, const double* c, double tau,
                   const void* Y, void* Z, int m_lo, int m_hi, long long* steps) {
    ADRT_REQUIRE(n >= 1 && n <= 11, "cheb_rows: need 1 <= n <= 11");
    ADRT_REQUIRE(K >= 0, "cheb_rows: K must be >= 0");
    ADRT_REQUIRE(1 <= m_lo && m_lo <= m_hi && m_hi <= n, "cheb_rows: need 1 <= m_lo <= m_hi <= n");
    if (K == 0) return OK;
    ADRT_REQUIRE(z && deg && a && c && Y && Z, "cheb_rows: null pointer");
    const int N = 1 << n;
    HostCallScratch& s = scratch();
    std::lock_guard<std::mutex> g(s.mu);
    ADRT_TRY(ensure_stream());
    const size_t rb = sizeof(c128) * (size_t)K * N;
    ADRT_TRY(ensure(0, sizeof(c128) * (size_t)K));
    ADRT_TRY(ensure(1, rb));
    ADRT_TRY(ensure(2, sizeof(int) * (size_t)K));
    ADRT_TRY(ensure(3, sizeof(double) * (size_t)K));
    ADRT_TRY(ensure(4, sizeof(double) * (size_t)K));
    ADRT_TRY(ensure(5, cheb_table_bytes(K, deg)));
    ADRT_HIP(hipMemcpyAsync(s.buf[0], z, sizeof(c128) * K, hipMemcpyHostToDevice, s.stream));
    ADRT_HIP(hipMemcpyAsync(s.buf[1], Y, rb, hipMemcpyHostToDevice, s.stream));
    ADRT_HIP(hipMemcpyAsync(s.buf[2], deg, sizeof(int) * K, hipMemcpyHostToDevice, s.stream));
    ADRT_HIP(hipMemcpyAsync(s.buf[3], a, sizeof(double) * K, hipMemcpyHostToDevice, s.stream));
    ADRT_HIP(hipMemcpyAsync(s.buf[4], c, sizeof(double) * K, hipMemcpyHostToDevice, s.stream));
    RowsArgs A{};
    A.Y = s.buf[1];
    A.Z = s.buf[1];
    A.rows = K;
    A.in_stride = A.out_stride = N;
    A.n = n;
    A.N = N;
    A.period = K;
    A.z = (const c128*)s.buf[0];
    A.deg = (const int*)s.buf[2];
    A.a = (const double*)s.buf[3];
    A.c = (const double*)s.buf[4];
    A.tau = tau;
    ADRT_TRY(cheb_table_fill(K, deg, (const int*)s.buf[2], (const double*)s.buf[3], (const double*)s.buf[4], s.buf[5],
                             s.stream, &A.wcoef, &A.wcoff));
    A.nops = 1;
    A.ops[0] = RowOp{OP_P, m_lo, m_hi};
    ADRT_TRY(launch_rows(A, PREC_F64, s.stream));
    ADRT_HIP(hipMemcpyAsync(Z, s.buf[1], rb, hipMemcpyDeviceToHost, s.stream));
    ADRT_HIP(hipStreamSynchronize(s.stream));
    if (steps) {
        long long tot = 0;
        for (int k = 0; k < K; ++k)
            if (deg[k] > 1) tot += (long long)(deg[k] - 1) * 2 * (m_hi - m_lo + 1);
        *steps += tot;
    }
    return OK;
}

// Certified bound by the row-sum recursion, one block per frequency; r lives in a global scratch of 2 (2A+1)
// doubles per frequency (cpp/oblique.cpp:108-130).
__global__ void cert_kernel(int n, int K, const double* omega, int m_lo, double* out, double* scr) {
    const long A = 1L << n;
    for (int k = blockIdx.x; k < K; k += gridDim.x) {
        double* r = scr + (size_t)k * 2 * (2 * A + 1);
        double* rn = r + (2 * A + 1);
        const double om = omega[k];
        for (long i = threadIdx.x; i < 2 * A + 1; i += blockDim.x) r[i] = 1.0;
        __syncthreads();
        for (int m = m_lo; m <= n; ++m) {
            const long amax = (1L << (n - m)) - 1;
            for (long i = threadIdx.x; i < 2 * A + 1; i += blockDim.x) rn[i] = 0.0;
            __syncthreads();
            for (long av = -amax + threadIdx.x; av <= amax; av += blockDim.x) {
                const double t0 = 2.0 * fabs(cos(0.5 * om * (double)av)) * r[2 * av + A];
                const double tp = 2.0 * fabs(cos(0.5 * om * (double)(av + 1))) * r[2 * av + 1 + A];
                const double tm = 2.0 * fabs(cos(0.5 * om * (double)(av - 1))) * r[2 * av - 1 + A];
                rn[av + A] = t0 + (tp > tm ? tp : tm);
            }
            __syncthreads();
            double* tmp = r;
            r = rn;
            rn = tmp;
        }
        if (threadIdx.x == 0) out[k] = r[A];
        __syncthreads();
    }
}

int cert_rowsum_impl(int n, int K, const double* omega, int m_lo, double* out) {
    ADRT_REQUIRE(n >= 1 && n <= 14, "cert_rowsum: need 1 <= n <= 14");
    ADRT_REQUIRE(K >= 0, "cert_rowsum: K must be >= 0");
    // level m reads r[2a +- 1 + A] for |a| <= 2^(n-m) - 1, inside the 2A + 1 slot only for m >= 1; m_lo > n runs no
    // level and gives 1, as the reference does
    ADRT_REQUIRE(m_lo >= 1, "cert_rowsum: need m_lo >= 1");
    if (K == 0) return OK;
    ADRT_REQUIRE(omega && out, "cert_rowsum: null pointer");
    HostCallScratch& s = scratch();
    std::lock_guard<std::mutex> g(s.mu);
    ADRT_TRY(ensure_stream());
    const long A = 1L << n;
    const int chunk = 4096;
    ADRT_TRY(ensure(0, sizeof(double) * (size_t)K));
    ADRT_TRY(ensure(1, sizeof(double) * (size_t)K));
    ADRT_TRY(ensure(2, sizeof(double) * 2 * (size_t)(2 * A + 1) * (K < chunk ? K : chunk)));
    ADRT_HIP(hipMemcpyAsync(s.buf[0], omega, sizeof(double) * K, hipMemcpyHostToDevice, s.stream));
    for (int k0 = 0; k0 < K; k0 += chunk) {
        const int kk = (K - k0) < chunk ? (K - k0) : chunk;
        hipLaunchKernelGGL(cert_kernel, dim3(kk), dim3(256), 0, s.stream, n, kk, (const double*)s.buf[0] + k0, m_lo,
                           (double*)s.buf[1] + k0, (double*)s.buf[2]);
        ADRT_TRY(launch_check("cert_kernel"));
    }
    ADRT_HIP(hipMemcpyAsync(out, s.buf[1], sizeof(double) * K, hipMemcpyDeviceToHost, s.stream));
    ADRT_HIP(hipStreamSynchronize(s.stream));
    return OK;
}

}  // namespace

// ---------------------------------------------------------------- device-resident row plan
struct RowsPlan {
    int n = 0, N = 0, L = 0, prec = PREC_F64, comp = 0;
    double tau = 0;
    c128* z = nullptr;
    int* deg = nullptr;
    double* a = nullptr;
    double* c = nullptr;
    int* dslot = nullptr;
    c128* dFT = nullptr;
    int n_dense = 0;
    void* ztab = nullptr;  // z_k^t per parameter row in the row type (ztab_fill)
    void* wtab = nullptr;  // Chebyshev step weights (cheb_table_fill)
    const c128* wcoef = nullptr;
    const int* wcoff = nullptr;
};

}  // namespace adrt

using namespace adrt;

extern "C" {

const char* adrt_last_error(void) { return last_error_cstr(); }

int adrt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int adrt_malloc(void** ptr, unsigned long long bytes) {
    ADRT_REQUIRE(ptr, "adrt_malloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? E_NOMEM : E_HIP;
    }
    return OK;
}

int adrt_free(void* ptr) {
    if (ptr) ADRT_HIP(hipFree(ptr));
    return OK;
}

int adrt_memcpy(void* dst, const void* src, unsigned long long bytes, int kind, void* stream) {
    if (bytes == 0) return OK;
    ADRT_REQUIRE(dst && src, "adrt_memcpy: null pointer");
    ADRT_REQUIRE(kind >= 0 && kind <= 4, "adrt_memcpy: kind must be 0..4 (h2h, h2d, d2h, d2d, d2h async)");
    const hipMemcpyKind k[5] = {hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                hipMemcpyDeviceToDevice, hipMemcpyDeviceToHost};
    ADRT_HIP(hipMemcpyAsync(dst, src, bytes, k[kind], (hipStream_t)stream));
    if (kind == 2 || kind == 0) ADRT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return OK;
}

int adrt_host_malloc(void** ptr, unsigned long long bytes) {
    ADRT_REQUIRE(ptr, "adrt_host_malloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return OK;
    const hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
        return E_NOMEM;
    }
    return OK;
}

int adrt_host_free(void* ptr) {
    if (ptr) ADRT_HIP(hipHostFree(ptr));
    return OK;
}

int adrt_memset(void* dst, int value, unsigned long long bytes, void* stream) {
    if (bytes == 0) return OK;
    ADRT_REQUIRE(dst, "adrt_memset: null pointer");
    ADRT_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return OK;
}

int adrt_stream_create(void** stream) {
    ADRT_REQUIRE(stream, "adrt_stream_create: null pointer");
    hipStream_t s;
    ADRT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return OK;
}

// Live own-queue streams of the process.  Each holds a hardware queue beside HIP's pool (GPU_MAX_HW_QUEUES, 4 by
// default), so their number is capped: a server needs one per request size it holds.
std::mutex g_own_mu;
std::vector<hipStream_t> g_own_queues;

int adrt_stream_create_own_queue(void** stream) {
    ADRT_REQUIRE(stream, "adrt_stream_create_own_queue: null pointer");
    std::lock_guard<std::mutex> g(g_own_mu);
    if ((int)g_own_queues.size() >= ADRT_MAX_OWN_QUEUES) {
        set_error("adrt_stream_create_own_queue: " + std::to_string(ADRT_MAX_OWN_QUEUES) +
                  " own-queue streams are alive (the cap); destroy one first or use adrt_stream_create");
        return E_STATE;
    }
    int dev = 0, cus = 0;
    ADRT_HIP(hipGetDevice(&dev));
    ADRT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    ADRT_REQUIRE(cus > 0 && cus <= 1024, "adrt_stream_create_own_queue: unexpected CU count");
    std::vector<uint32_t> mask((cus + 31) / 32, 0u);
    for (int c = 0; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
    hipStream_t s;
    // a stream with a CU mask (here: every CU) is given a hardware queue of its own instead of one from the shared pool
    ADRT_HIP(hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()));
    g_own_queues.push_back(s);
    *stream = (void*)s;
    return OK;
}

int adrt_own_queue_count(void) {
    std::lock_guard<std::mutex> g(g_own_mu);
    return (int)g_own_queues.size();
}

int adrt_stream_destroy(void* stream) {
    if (!stream) return OK;
    {
        std::lock_guard<std::mutex> g(g_own_mu);
        for (size_t i = 0; i < g_own_queues.size(); ++i)
            if (g_own_queues[i] == (hipStream_t)stream) {
                g_own_queues.erase(g_own_queues.begin() + i);
                break;
            }
    }
    ADRT_HIP(hipStreamDestroy((hipStream_t)stream));
    return OK;
}

int adrt_stream_sync(void* stream) {
    ADRT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return OK;
}

int adrt_device_sync(void) {
    ADRT_HIP(hipDeviceSynchronize());
    return OK;
}

int adrt_device_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes) {
    size_t f = 0, t = 0;
    ADRT_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return OK;
}

// Reference-compatible entry points (void, like cpp/oblique.cpp); failures print and set adrt_last_error().
void bfly_rows(int n, int K, const adrt_c128* z, adrt_c128* X, int m_lo, int m_hi, int transpose, int nthreads) {
    (void)nthreads;
    int st = bfly_rows_impl(n, K, z, X, m_lo, m_hi, transpose);
    report(st, "bfly_rows");
}

void cheb_rows(int n, int K, const adrt_c128* z, const int* deg, const double* a, const double* c, double tau,
               const adrt_c128* Y, adrt_c128* Z, int m_lo, int m_hi, int nthreads, long long* steps) {
    (void)nthreads;
    int st = cheb_rows_impl(n, K, z, deg, a, c, tau, Y, Z, m_lo, m_hi, steps);
    report(st, "cheb_rows");
}

void cert_rowsum(int n, int K, const double* omega, int m_lo, double* out, int nthreads) {
    (void)nthreads;
    int st = cert_rowsum_impl(n, K, omega, m_lo, out);
    report(st, "cert_rowsum");
}

int omp_max_threads(void) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
    return cus > 0 ? cus : 1;
}

int adrt_rows_config_set(int n, int precision, int loge) { return rows_config_set(n, precision, loge); }
int adrt_rows_config_get(int n, int precision) { return rows_config_get(n, precision); }
int adrt_dense_config_set(int n, int precision, int mode) { return dense_config_set(n, precision, mode); }
int adrt_dense_config_get(int n, int precision) { return dense_config_get(n, precision); }

int adrt_bfly_rows(int n, int K, const adrt_c128* z, adrt_c128* X, int m_lo, int m_hi, int transpose) {
    return bfly_rows_impl(n, K, z, X, m_lo, m_hi, transpose);
}

int adrt_cheb_rows(int n, int K, const adrt_c128* z, const int* deg, const double* a, const double* c, double tau,
                   const adrt_c128* Y, adrt_c128* Z, int m_lo, int m_hi, long long* steps) {
    return cheb_rows_impl(n, K, z, deg, a, c, tau, Y, Z, m_lo, m_hi, steps);
}

int adrt_cert_rowsum(int n, int K, const double* omega, int m_lo, double* out) {
    return cert_rowsum_impl(n, K, omega, m_lo, out);
}

int adrt_rows_plan_create(adrt_rows_plan** out, int n, int L, const adrt_c128* z, const int* deg, const double* a,
                          const double* c, double tau, int n_dense, const int* dense_k, const adrt_c128* dense_F,
                          int precision) {
    ADRT_REQUIRE(out, "rows_plan_create: null out pointer");
    *out = nullptr;
    ADRT_REQUIRE(n >= 1 && n <= 11, "rows_plan_create: need 1 <= n <= 11");
    ADRT_REQUIRE(L >= 1, "rows_plan_create: L must be positive");
    ADRT_REQUIRE(z && deg && a && c, "rows_plan_create: null parameter array");
    ADRT_REQUIRE(n_dense >= 0 && n_dense <= L, "rows_plan_create: bad n_dense");
    ADRT_REQUIRE(n_dense == 0 || (dense_k && dense_F), "rows_plan_create: dense rows need dense_k and dense_F");
    ADRT_REQUIRE(precision >= 0 && precision <= 2, "rows_plan_create: precision must be 0, 1 or 2");
    auto* p = new RowsPlan();
    p->n = n;
    p->N = 1 << n;
    p->L = L;
    p->prec = precision == PREC_F64 ? PREC_F64 : PREC_F32;
    p->comp = precision == 1;
    p->tau = tau;
    p->n_dense = n_dense;
    auto fail = [&](int st) {
        adrt_rows_plan_destroy((adrt_rows_plan*)p);
        return st;
    };
    const int N = p->N;
    int st;
    if ((st = adrt_malloc((void**)&p->z, sizeof(c128) * L)) != OK) return fail(st);
    if ((st = adrt_malloc((void**)&p->deg, sizeof(int) * L)) != OK) return fail(st);
    if ((st = adrt_malloc((void**)&p->a, sizeof(double) * L)) != OK) return fail(st);
    if ((st = adrt_malloc((void**)&p->c, sizeof(double) * L)) != OK) return fail(st);
    if (hipMemcpy(p->z, z, sizeof(c128) * L, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->deg, deg, sizeof(int) * L, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->a, a, sizeof(double) * L, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->c, c, sizeof(double) * L, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("rows_plan_create: upload failed");
        return fail(E_HIP);
    }
    if ((st = adrt_malloc(&p->wtab, cheb_table_bytes(L, deg))) != OK ||
        (st = cheb_table_fill(L, deg, p->deg, p->a, p->c, p->wtab, nullptr, &p->wcoef, &p->wcoff)) != OK ||
        (st = adrt_malloc(&p->ztab, ztab_bytes(n, L, p->prec))) != OK ||
        (st = ztab_fill(n, L, p->z, p->prec, p->ztab, nullptr)) != OK)
        return fail(st);
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        set_error("rows_plan_create: Chebyshev weight table failed");
        return fail(E_HIP);
    }
    if (n_dense > 0) {
        std::vector<int> slot(L, -1);
        for (int i = 0; i < n_dense; ++i) {
            if (dense_k[i] < 0 || dense_k[i] >= L) {
                set_error("rows_plan_create: dense_k out of range");
                return fail(E_INVALID);
            }
            slot[dense_k[i]] = i;
        }
        // store F^T so the kernel's thread i streams column i of F with coalesced reads
        std::vector<c128> ft((size_t)n_dense * N * N);
        const c128* F = (const c128*)dense_F;
        for (int s = 0; s < n_dense; ++s)
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) ft[(size_t)s * N * N + (size_t)j * N + i] = F[(size_t)s * N * N + (size_t)i * N + j];
        if ((st = adrt_malloc((void**)&p->dslot, sizeof(int) * L)) != OK) return fail(st);
        if ((st = adrt_malloc((void**)&p->dFT, sizeof(c128) * ft.size())) != OK) return fail(st);
        if (hipMemcpy(p->dslot, slot.data(), sizeof(int) * L, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->dFT, ft.data(), sizeof(c128) * ft.size(), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("rows_plan_create: dense upload failed");
            return fail(E_HIP);
        }
    }
    *out = (adrt_rows_plan*)p;
    return OK;
}

int adrt_rows_plan_destroy(adrt_rows_plan* plan) {
    auto* p = (RowsPlan*)plan;
    if (!p) return OK;
    (void)hipFree(p->z);
    (void)hipFree(p->deg);
    (void)hipFree(p->a);
    (void)hipFree(p->c);
    (void)hipFree(p->dslot);
    (void)hipFree(p->dFT);
    (void)hipFree(p->wtab);
    (void)hipFree(p->ztab);
    delete p;
    return OK;
}

int adrt_rows_run(adrt_rows_plan* plan, const void* dY, void* dZ, long long rows, const int* ops, int nops,
                  void* stream) {
    auto* p = (RowsPlan*)plan;
    ADRT_REQUIRE(p, "rows_run: null plan");
    ADRT_REQUIRE(rows >= 0, "rows_run: rows must be >= 0");
    ADRT_REQUIRE(ops || nops == 0, "rows_run: null op list");
    ADRT_REQUIRE(nops >= 0 && nops <= MAX_OPS, "rows_run: at most 8 ops");
    if (rows == 0) return OK;
    ADRT_REQUIRE(dY && dZ, "rows_run: null row buffer");
    RowsArgs A{};
    A.Y = dY;
    A.Z = dZ;
    A.rows = rows;
    A.in_stride = A.out_stride = p->N;
    A.n = p->n;
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
    A.ztab = p->ztab;
    A.nops = nops;
    for (int i = 0; i < nops; ++i) A.ops[i] = RowOp{ops[3 * i], ops[3 * i + 1], ops[3 * i + 2]};
    A.comp = p->comp;
    return launch_rows(A, p->prec, (hipStream_t)stream);
}

}  // extern "C"
