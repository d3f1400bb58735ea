// Packed-pass kernels around the row engine (offset FFTs, quadrant packing / unpacking) and the image-space
// kernels of the outer iteration (G_1 stencil, lorr projector, Chebyshev updates).
#pragma once
#include "common.hpp"

namespace adrt {

// Forward FFT over offsets of the packed quadrant strips (fastops.py:209-214).
//   mode 0 (image):   strip_q = T_q x zero-padded to L rows (fastspec.py:142-152)
//   mode 1 (data):    strip_q = iota(b_q): offset h in [-(N-1), N) -> row h mod L (fastspec.py:170-174)
// Output rows: Y[((s*2 + pair)*L + k)*N + col], pair 0 = q0 + i q1, pair 1 = q2 + i q3.
int launch_pack_fft(int n, int mode, const double* src, void* Y, int S, int prec, const void* tw, hipStream_t st);

// Inverse FFT over offsets and split of the packed pairs into the four real quadrant strips, rows h < N only
// (the window keeps nothing else, fastops.py:216-220): Sq[((s*4 + q)*N + h)*N + col].
int launch_ifft_unpack(int n, const void* Z, void* Sq, int S, int prec, const void* tw, hipStream_t st);

// out[s] = sum_q TqT(q, Sq[s, q]) windowed to N x N (fastspec.py:154-168), summed as (q0 + q1) + (q2 + q3).
// If y != nullptr the kernel writes y - out instead (the outer residual).  out is double, or float with out_f32 (the
// residual of an FP32 outer step, which the float G_1 iteration rounds to float anyway).
int launch_combine(int n, const void* Sq, int prec, const double* y, void* out, int S, hipStream_t st,
                   bool out_f32 = false);

// Two-step register FFT for 6 <= n <= 9 (fft2.hip); mode 0 / 1 as launch_pack_fft, mode 2 as launch_ifft_unpack.
bool fft2_supported(int n);
int launch_fft2(int n, int mode, const void* src, void* dst, int S, int prec, const void* tw, hipStream_t st);
// The inverse FFT with the four strips summed in the two pairs that need no transpose between them (fft2.hip):
//   UV[s, 0] = U, U[h][a] = S0[h][a] + S3[h][N-1-a];   UV[s, 1] = V, V[a][c] = S1[N-1-a][c] + S2[a][c]
// (each strip value rounded to float first, as launch_ifft_unpack stores it, and the sum rounded to float): half the
// strip traffic of launch_ifft_unpack + launch_combine.  A block runs four column FFTs in turn, so it has a quarter of
// launch_ifft_unpack's blocks: ifft_pair_blocks(n, prec) per sinogram, 0 where unsupported (float, 7 <= n <= 9).
// launch_combine_pair then adds U^T + V.
int ifft_pair_blocks(int n, int prec);
int launch_ifft_pair(int n, const void* Z, void* UV, int S, int prec, const void* tw, hipStream_t st);
// out[s] = U^T + V windowed to N x N: out[a][c] = U[N-1-c][a] + V[a][c], or y - that; types as launch_combine.
int launch_combine_pair(int n, const void* UV, int prec, const double* y, void* out, int S, hipStream_t st,
                        bool out_f32 = false);
// Three-step register FFT for n = 10, 11 (fft3.hip), modes as launch_fft2.
bool fft3_supported(int n);
int launch_fft3(int n, int mode, const void* src, void* dst, int S, int prec, const void* tw, hipStream_t st);

// Twiddle table exp(sign * 2 pi i m / L), m < L/2, in the precision's type (host builds it once per plan).
size_t twiddle_bytes(int n, int prec);

// Image-space kernels on S images of N x N doubles.
int launch_scale(const double* x, double alpha, double* out, long long count, hipStream_t st);
// out = x + w1 * (a * r) + w0 * (x - xp)          (fastops.py:255-256; fasthybrid.py:123-124); xp == null means 0
int launch_cheb_update(const double* x, const double* xp, const double* r, double a, double w1, double w0,
                       double* out, long long count, hipStream_t st);
// out = x + w1 * (a * (y - G1 x)) + w0 * (x - xp) (one step of g1_inv, fasthybrid.py:95-98, 114-126); xp == null
// means the zero start of the first step.
int launch_g1_step(int n, const double* y, const double* x, const double* xp, double a, double w1, double w0,
                   double* out, int S, hipStream_t st);
// Per-image weights of the mode-2 epilogue below, for a batch whose images sit at different outer steps
// (adrt_plan_tick): image s takes w1[s], w0[s], and its x_prev only where bit s of has_xp is set (else 0).
constexpr int G1_MAX_LANES = 64;
struct G1Lanes {
    double w1[G1_MAX_LANES], w0[G1_MAX_LANES];
    unsigned long long has_xp = 0;
};
// g1_inv(y) (deg Chebyshev steps of G_1 on [4, 16]) in one launch, with an epilogue:
//   mode 0: out = P y;  mode 1: out = a_out * P y;  mode 2: out = x + w1 * (a_out * P y) + w0 * (x - xp) (xp null = 0)
// prec PREC_F32 runs the inner iteration in float (outer iterates and the epilogue stay double); its right-hand side
// may then come as float (y32, with y null).  s_tile (0: S) is the batch size that picks the double iteration's tile
// size; a plan passes its max_batch, so that an image's bits do not depend on how many others share the launch.
// lanes (S <= G1_MAX_LANES) replaces w1, w0 and the use of xp per image.
int launch_g1_fused(int n, int deg, int mode, const double* y, const double* x, const double* xp, double a_out,
                    double w1, double w0, double* out, int S, hipStream_t st, int prec = PREC_F64,
                    const float* y32 = nullptr, int s_tile = 0, const G1Lanes* lanes = nullptr);
// out = (y - Pi y)/4 + Pi y / 12 (fasthybrid.py:101-111)
int launch_lorr(int n, const double* y, double* out, int S, hipStream_t st);

// Forward / transpose ADRT (adrtcore.py:63-137).  dtype 0: int32 (transpose accumulates into int64),
// 1: float, 2: double.  work holds one (S, 4, 2N-1, N) array of the intermediate type.
int launch_adrt_forward(int n, int dtype, const void* x, void* b, void* work, int S, hipStream_t st);
int launch_adrt_transpose(int n, int dtype, const void* b, void* x, void* work, int S, hipStream_t st);
size_t adrt_work_bytes(int n, int dtype, int S, int transpose);

}  // namespace adrt
