// qp_internal.h -- shared host/device definitions of the MI355X slice loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/quetzal_hip.h"

namespace qp {

// ----------------------------------------------------------------- errors
void set_error(const std::string &msg);
#define QP_HIP_CHECK(expr)                                                                    \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::qp::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));              \
            return QP_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)
#define QP_TRY(expr)                     \
    do {                                 \
        int _s = (expr);                 \
        if (_s < 0) return _s;           \
    } while (0)

// ------------------------------------------------------------ FFT plan
constexpr int kMaxStages = 24;
constexpr int kMaxN = 1024;          // longest transform the LDS kernels take
constexpr int kLdsComplex = 1024;    // complex doubles per LDS data buffer (16 KiB; two buffers)

// One transform axis as the kernels see it (passed by value).
struct FftAxisDev {
    int n;
    int nstages;
    int radix[kMaxStages];
    const double2 *tw;  // e^{-2 pi i t / n}, t in [0, n)
};

struct FftAxis {
    FftAxisDev dev{};
    double2 *d_tw = nullptr;
    int plan(int n);   // radix plan only (host); dev.tw left null
    int init(int n);   // plan + twiddle table in device memory
    void destroy();
};

// 2D periodic solver geometry (src/System_3d.cpp:686-706).
struct Solver2d {
    int n1 = 0, n2 = 0, nh = 0;
    double d1 = 0, d2 = 0;
    FftAxis ax1, ax2;
    double *d_lam1 = nullptr;  // (2cos(2 pi j/n1) - 2)/d1/d1
    double *d_lam2 = nullptr;  // (2cos(2 pi k/n2) - 2)/d2/d2
    double2 *spec[2] = {nullptr, nullptr};  // n1 x nh half spectra (two for paired solves), nbatch each
    int nbatch = 1;                           // wakes sharing this solver (spectra side by side)
    int rows_per_block = 1;   // complex row transforms (= 2 real rows) per block
    int cols_per_block = 1;
    // 1: both sides 2^p in [16, 1024] -> register-resident FFT kernels with the
    // column-major spectrum spec[l * n1 + j]; 0: generic kernels, row-major
    // spec[j * nh + l] (FFTW's r2c layout).  Fourier2d always uses the latter.
    int wave = 0;
    int init(int n1, int n2, double d1, double d2, int nbatch = 1);
    void default_shape();  // wave / rows_per_block / cols_per_block from n1, n2
    int set_steps(double d1, double d2);
    int autotune(hipStream_t s, double *scratch_real);
    void destroy();
};

// ---------------------------------------------------------- control block
// Soft stop of src/System_3d.cpp:541-591 without host round trips:
//  stop_acc: set by the B store epilogue when |B| > threshold or NaN;
//  stop_now: latched by the first kernel after the iteration that set stop_acc.
// B-solve kernels skip when stop_now is set, every other slice kernel skips
// when stop_acc is set, so the iteration that trips the rule completes (the
// reference finishes both B solves before breaking, :582-584).
struct Ctl {
    int stop_acc;
    int stop_now;
    int stop_slice;
    int slice_of[2];   // mixed path: slice index for the even / odd slice step
    unsigned ovf;      // mixed path: particle blocks whose deposit window exceeded the LDS capacity
    int pad[2];
    double border[4];  // plain path: border mean of the last psi solve
};

__device__ __forceinline__ bool ctl_skip(const Ctl *c) { return c->stop_acc != 0; }
__device__ __forceinline__ bool ctl_skip_b(const Ctl *c) { return c->stop_now != 0; }
// c is the control block of the launching block's wake.  Every skipping block
// latches (the same value): in a group launch blockIdx.x is a flat index over
// all batches and members, so no single block index is present for each wake.
__device__ __forceinline__ void ctl_latch(Ctl *c) {
    if (threadIdx.x == 0) c->stop_now = 1;
}

// ------------------------------------------------------------ complex math
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ------------------------------------------------------- CIC index math
// Memory-safety guard: an index outside [0, n) (only reachable from a
// non-finite coordinate) is sent to 0 instead of addressing outside the grid.
__device__ __forceinline__ void idx_guard(int &j, int n) {
    if ((unsigned)j >= (unsigned)n) j = 0;
}

// Node-origin CIC index, src/array_utils.cpp:90-103: j1 = floor((y-o)/d),
// j2 = (j1+1) % n, frac = (y-o)/d - j1, then a negative j1 wraps.  An index
// of exactly n (y/d rounded up to n) wraps to 0 (the reference's asserts are
// compiled out there).
__device__ __forceinline__ void cic_floor(double y, double d, double o, int n, int &j1, int &j2, double &f) {
    const double t = (y - o) / d;
    j1 = (int)floor(t);
    j2 = (j1 + 1) % n;
    f = t - j1;
    if (j1 < 0) j1 += n;
    if (j1 >= n) j1 -= n;
    idx_guard(j1, n);
    idx_guard(j2, n);
}
// Truncating variant without origin, src/array_utils.cpp:121-127.
__device__ __forceinline__ void cic_trunc(double y, double d, int n, int &j1, int &j2, double &f) {
    const double t = y / d;
    j1 = (int)t;
    j2 = (j1 + 1) % n;
    f = t - j1;
    if (j1 >= n) j1 -= n;
    idx_guard(j1, n);
    idx_guard(j2, n);
}

__device__ __forceinline__ double gather4(const double *a, int n2, int j1, int j2, int k1, int k2, double yf,
                                          double zf) {
    return a[(size_t)n2 * j1 + k1] * (1 - yf) * (1 - zf) + a[(size_t)n2 * j2 + k1] * yf * (1 - zf) +
           a[(size_t)n2 * j1 + k2] * (1 - yf) * zf + a[(size_t)n2 * j2 + k2] * yf * zf;
}

__device__ __forceinline__ void scatter4_global(double *a, int n2, int j1, int j2, int k1, int k2, double yf,
                                                double zf, double v) {
    unsafeAtomicAdd(&a[(size_t)n2 * j1 + k1], v * (1 - yf) * (1 - zf));
    unsafeAtomicAdd(&a[(size_t)n2 * j2 + k1], v * yf * (1 - zf));
    unsafeAtomicAdd(&a[(size_t)n2 * j1 + k2], v * (1 - yf) * zf);
    unsafeAtomicAdd(&a[(size_t)n2 * j2 + k2], v * yf * zf);
}

// array_to_particle(y,z,array2d&), src/array_utils.cpp:167-188.
__device__ __forceinline__ double gather_o(double y, double z, const double *a, int n1, int n2, double d1,
                                           double d2, double o1, double o2) {
    int j1, j2, k1, k2;
    double yf, zf;
    cic_floor(y, d1, o1, n1, j1, j2, yf);
    cic_floor(z, d2, o2, n2, k1, k2, zf);
    return gather4(a, n2, j1, j2, k1, k2, yf, zf);
}

// array_yder_to_particle(y,z,array2d&), src/array_utils.cpp:240-263.
__device__ __forceinline__ double gather_yder(double y, double z, const double *a, int n1, int n2, double d1,
                                              double d2) {
    int j1 = (int)floor(y / d1 - 0.5);
    const double yf = (y / d1 - 0.5) - j1;
    if (j1 < 0) j1 += n1;
    idx_guard(j1, n1);
    const int j2 = (j1 + 1) % n1, j3 = (j1 + 2) % n1;
    int k1, k2;
    double zf;
    cic_trunc(z, d2, n2, k1, k2, zf);
    const double g11 = (a[(size_t)n2 * j2 + k1] - a[(size_t)n2 * j1 + k1]) / d1;
    const double g21 = (a[(size_t)n2 * j3 + k1] - a[(size_t)n2 * j2 + k1]) / d1;
    const double g12 = (a[(size_t)n2 * j2 + k2] - a[(size_t)n2 * j1 + k2]) / d1;
    const double g22 = (a[(size_t)n2 * j3 + k2] - a[(size_t)n2 * j2 + k2]) / d1;
    return g11 * (1 - yf) * (1 - zf) + g21 * yf * (1 - zf) + g12 * (1 - yf) * zf + g22 * yf * zf;
}

// array_zder_to_particle(y,z,array2d&), src/array_utils.cpp:265-288.
__device__ __forceinline__ double gather_zder(double y, double z, const double *a, int n1, int n2, double d1,
                                              double d2) {
    int j1, j2;
    double yf;
    cic_trunc(y, d1, n1, j1, j2, yf);
    int k1 = (int)floor(z / d2 - 0.5);
    const double zf = (z / d2 - 0.5) - k1;
    if (k1 < 0) k1 += n2;
    idx_guard(k1, n2);
    const int k2 = (k1 + 1) % n2, k3 = (k1 + 2) % n2;
    const double g11 = (a[(size_t)n2 * j1 + k2] - a[(size_t)n2 * j1 + k1]) / d2;
    const double g21 = (a[(size_t)n2 * j2 + k2] - a[(size_t)n2 * j2 + k1]) / d2;
    const double g12 = (a[(size_t)n2 * j1 + k3] - a[(size_t)n2 * j1 + k2]) / d2;
    const double g22 = (a[(size_t)n2 * j2 + k3] - a[(size_t)n2 * j2 + k2]) / d2;
    return g11 * (1 - yf) * (1 - zf) + g21 * yf * (1 - zf) + g12 * (1 - yf) * zf + g22 * yf * zf;
}

// normalize_coordinates, src/System_3d.cpp:668-684.
__device__ __forceinline__ void wrap_coords(double &y, double &z, double ly, double lz) {
    y = fmod(y, ly);
    if (y < 0) y += ly;
    z = fmod(z, lz);
    if (z < 0) z += lz;
}

// ----------------------------------------------------- solver launches
// All launches go to `s`.  `src_fn` kinds select what the row-forward pass reads.
int solve_rows_fwd_plain(const Solver2d &S, const double *src, double2 *spec, hipStream_t s);
int solve_cols(const Solver2d &S, double D, int nspec, hipStream_t s, const Ctl *ctl, int skip_mode);
// Column pass of nw wakes side by side (spectra at + wake * n1 * nh, control blocks at ctl + wake).
int solve_cols_batch(const Solver2d &S, double D, int nspec, int nw, hipStream_t s, const Ctl *ctl,
                     int skip_mode);
int solve_rows_inv_plain(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s);
// Plain-path B solve epilogue: store v/n1/n2 and flag |B| > thr or NaN.
int solve_rows_inv_bcheck(const Solver2d &S, const double2 *spec, double *dst, double thr, Ctl *ctl,
                          int slice, hipStream_t s);
size_t rows_lds(const Solver2d &S);
size_t cols_lds(const Solver2d &S);
int rows_blocks(const Solver2d &S);
int solve_cols_forward_only(const Solver2d &S, hipStream_t s);
int solve_cols_inverse_only(const Solver2d &S, hipStream_t s);
// Generic kernels only: column transforms of nspec spectra, mode 0 forward, 1 inverse.
int solve_cols_pass(const Solver2d &S, int nspec, int mode, hipStream_t s);
int solve_rows_inv_unscaled(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s);
// Fourier2d transforms in FFTW's layout (generic kernels whatever S.wave says).
int fourier_rows_fwd(const Solver2d &S, const double *src, double2 *spec, hipStream_t s);
int fourier_rows_inv(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s);
int solve_rows_fwd_plain_ctl(const Solver2d &S, const double *src, double2 *spec, hipStream_t s,
                             const Ctl *ctl, int skip_mode);
int solve_rows_inv_plain_ctl(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s,
                             const Ctl *ctl);

}  // namespace qp

namespace qp {
// Per-GPU set-up of the allocation-free solve (qp_noalloc.hip): resolves the
// device address of its file-scope scratch array on `device` (current).
int noalloc_setup(int device);
}  // namespace qp

struct qp_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
};
