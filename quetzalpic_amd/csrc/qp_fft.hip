// qp_fft.hip -- periodic 2D solve on the device, FP64.
//
// Replaces Fourier2d (src/Fourier2d.h:4-19) and
// System_3d::solve_poisson_equation (src/System_3d.cpp:686-706):
//   r2c over n1 x n2 -> (0,0) := 0 -> divide by the 5-point eigenvalue
//   -D + (2cos(2 pi j/n1) - 2)/d1^2 + (2cos(2 pi k/n2) - 2)/d2^2 -> c2r.
// Three launches: rows forward (two real rows packed per complex transform),
// columns forward + divide + inverse (one block owns whole columns), rows
// inverse (Hermitian rows packed two per complex transform).  Power-of-two
// grids from 16 to 1024 per side use the register-resident kernels of
// qp_wave_solver.h; other sizes, and Fourier2d's transforms (FFTW's r2c
// layout), use the generic LDS kernels of qp_solver_kernels.h.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "qp_fft_lds.h"
#include "qp_solver_kernels.h"
#include "qp_wave_solver.h"

namespace qp {

int FftAxis::plan(int n) {
    if (n < 1 || n > kMaxN) {
        set_error("transform length " + std::to_string(n) + " outside [1, " + std::to_string(kMaxN) + "]");
        return QP_ERR_ARG;
    }
    dev.n = n;
    dev.nstages = 0;
    int m = n;
    auto push = [&](int r) -> bool {
        if (dev.nstages >= kMaxStages) return false;
        dev.radix[dev.nstages++] = r;
        return true;
    };
    while (m % 4 == 0) { if (!push(4)) return QP_ERR_ARG; m /= 4; }
    while (m % 2 == 0) { if (!push(2)) return QP_ERR_ARG; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { if (!push(f)) return QP_ERR_ARG; m /= f; }
        if ((long)f * f > m && m > 1) { if (!push(m)) return QP_ERR_ARG; m = 1; }
    }
    dev.tw = nullptr;
    return QP_OK;
}

int FftAxis::init(int n) {
    QP_TRY(plan(n));
    std::vector<double2> tw(n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
        long double a = two_pi * (long double)t / (long double)n;
        tw[t] = make_double2((double)cosl(a), (double)(-sinl(a)));
    }
    QP_HIP_CHECK(hipMalloc(&d_tw, sizeof(double2) * n));
    QP_HIP_CHECK(hipMemcpy(d_tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    dev.tw = d_tw;
    return QP_OK;
}

void FftAxis::destroy() {
    if (d_tw) (void)hipFree(d_tw);
    d_tw = nullptr;
}

int Solver2d::set_steps(double d1_, double d2_) {
    d1 = d1_;
    d2 = d2_;
    // Same expression as the reference (src/System_3d.cpp:699), split per axis.
    std::vector<double> l1(n1), l2(nh);
    for (int j = 0; j < n1; ++j) l1[j] = (2 * cos(2 * M_PI * j / n1) - 2) / d1 / d1;
    for (int k = 0; k < nh; ++k) l2[k] = (2 * cos(2 * M_PI * k / n2) - 2) / d2 / d2;
    QP_HIP_CHECK(hipMemcpy(d_lam1, l1.data(), sizeof(double) * n1, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(d_lam2, l2.data(), sizeof(double) * nh, hipMemcpyHostToDevice));
    return QP_OK;
}

int Solver2d::init(int n1_, int n2_, double d1_, double d2_, int nbatch_) {
    nbatch = nbatch_ < 1 ? 1 : nbatch_;
    n1 = n1_;
    n2 = n2_;
    nh = n2 / 2 + 1;
    QP_TRY(ax1.init(n1));
    QP_TRY(ax2.init(n2));
    QP_HIP_CHECK(hipMalloc(&d_lam1, sizeof(double) * n1));
    QP_HIP_CHECK(hipMalloc(&d_lam2, sizeof(double) * nh));
    QP_TRY(set_steps(d1_, d2_));
    for (int i = 0; i < 2; ++i) QP_HIP_CHECK(hipMalloc(&spec[i], sizeof(double2) * (size_t)n1 * nh * nbatch));
    default_shape();
    return QP_OK;
}

// Kernel family and launch shape of an n1 x n2 solver (no device work).
void Solver2d::default_shape() {
    // QP_FFT=generic keeps the generic kernels for every size (A/B comparisons).
    const char *sel = getenv("QP_FFT");
    wave = (wf_supported(n1) && wf_supported(n2) && !(sel && strcmp(sel, "generic") == 0)) ? 1 : 0;
    const int npairs = (n1 + 1) / 2;
    int b = npairs / 64;
    if (b < 1) b = 1;
    if (b > kLdsComplex / n2) b = kLdsComplex / n2;
    rows_per_block = b < 1 ? 1 : b;
    int c = nh / 64;
    if (c < 1) c = 1;
    if (c > kLdsComplex / n1) c = kLdsComplex / n1;
    cols_per_block = c < 1 ? 1 : c;
}

// Pick rows_per_block / cols_per_block for this grid by timing the row and
// column passes (20 launches each) on scratch data; cached per (n1, n2).
int Solver2d::autotune(hipStream_t s, double *scratch_real) {
    if (wave) return QP_OK;  // the register-resident kernels have a fixed shape
    // shared by every solver of the process; wakes may be created from several
    // host threads, so lookups and the tuning that fills an entry hold the lock
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::pair<int, int>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(n1, n2);
    auto it = cache.find(key);
    if (it != cache.end()) {
        rows_per_block = it->second.first;
        cols_per_block = it->second.second;
        return QP_OK;
    }
    hipEvent_t e0, e1;
    QP_HIP_CHECK(hipEventCreate(&e0));
    QP_HIP_CHECK(hipEventCreate(&e1));
    QP_HIP_CHECK(hipMemsetAsync(scratch_real, 0, sizeof(double) * (size_t)n1 * n2, s));
    float best = 1e30f;
    int best_b = rows_per_block;
    for (int b = 1; b <= 16 && b * n2 <= kLdsComplex; b *= 2) {
        if (b > 1 && (n1 + 1) / 2 < b) break;
        rows_per_block = b;
        for (int w = 0; w < 3; ++w) QP_TRY(solve_rows_fwd_plain(*this, scratch_real, spec[0], s));
        QP_HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < 20; ++r) {
            QP_TRY(solve_rows_fwd_plain(*this, scratch_real, spec[0], s));
            QP_TRY(solve_rows_inv_unscaled(*this, spec[0], scratch_real, s));
        }
        QP_HIP_CHECK(hipEventRecord(e1, s));
        QP_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        QP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_b = b;
        }
    }
    rows_per_block = best_b;
    best = 1e30f;
    int best_c = cols_per_block;
    for (int c = 1; c <= 16 && c * n1 <= kLdsComplex; c *= 2) {
        if (c > 1 && nh < c) break;
        cols_per_block = c;
        for (int w = 0; w < 3; ++w) QP_TRY(solve_cols(*this, 10.0, 1, s, nullptr, 0));
        QP_HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < 20; ++r) QP_TRY(solve_cols(*this, 10.0, 1, s, nullptr, 0));
        QP_HIP_CHECK(hipEventRecord(e1, s));
        QP_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        QP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_c = c;
        }
    }
    cols_per_block = best_c;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cache[key] = std::make_pair(rows_per_block, cols_per_block);
    return QP_OK;
}

void Solver2d::destroy() {
    ax1.destroy();
    ax2.destroy();
    if (d_lam1) (void)hipFree(d_lam1);
    if (d_lam2) (void)hipFree(d_lam2);
    for (int i = 0; i < 2; ++i)
        if (spec[i]) (void)hipFree(spec[i]);
    d_lam1 = d_lam2 = nullptr;
    spec[0] = spec[1] = nullptr;
}

// ------------------------------------------------------------ launches
// LDS of a launch: two data buffers, the twiddle table (and for the column
// pass the n1 eigenvalue terms).
size_t rows_lds(const Solver2d &S) {
    return 2 * sizeof(double2) * (size_t)S.rows_per_block * S.n2 + sizeof(double2) * (size_t)S.n2;
}
size_t cols_lds(const Solver2d &S) {
    return 2 * sizeof(double2) * (size_t)S.cols_per_block * S.n1 + sizeof(double2) * (size_t)S.n1 +
           sizeof(double) * (size_t)S.n1;
}
int rows_blocks(const Solver2d &S) {
    const int npairs = (S.n1 + 1) / 2;
    return (npairs + S.rows_per_block - 1) / S.rows_per_block;
}

static int gen_rows_fwd(const Solver2d &S, const double *src, double2 *spec, hipStream_t s, const Ctl *ctl,
                        int skip_mode) {
    SrcPlain f{src, S.n2};
    hipLaunchKernelGGL(k_rows_fwd<SrcPlain>, dim3(rows_blocks(S)), dim3(kSolverThreads), rows_lds(S), s, f,
                       spec, spec, S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, ctl, skip_mode, 1, (size_t)0);
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

int solve_rows_fwd_plain_ctl(const Solver2d &S, const double *src, double2 *spec, hipStream_t s,
                             const Ctl *ctl, int skip_mode) {
    if (S.wave) return wlaunch_rows_fwd(S, SrcPlain{src, S.n2}, spec, spec, 1, s, ctl, skip_mode);
    return gen_rows_fwd(S, src, spec, s, ctl, skip_mode);
}

int solve_rows_fwd_plain(const Solver2d &S, const double *src, double2 *spec, hipStream_t s) {
    return solve_rows_fwd_plain_ctl(S, src, spec, s, nullptr, 0);
}

int fourier_rows_fwd(const Solver2d &S, const double *src, double2 *spec, hipStream_t s) {
    return gen_rows_fwd(S, src, spec, s, nullptr, 0);
}

static int cols_launch(const Solver2d &S, double D, int nspec, hipStream_t s, const Ctl *ctl, int skip_mode,
                       int mode) {
    const int nb = (S.nh + S.cols_per_block - 1) / S.cols_per_block;
    hipLaunchKernelGGL(k_cols_solve, dim3(nb, nspec), dim3(kSolverThreads), cols_lds(S), s, S.spec[0], S.spec[1],
                       S.n1, S.nh, S.ax1.dev, (const double *)S.d_lam1, (const double *)S.d_lam2, D,
                       S.cols_per_block, ctl, skip_mode, mode, nspec, (size_t)0);
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

int solve_cols_batch(const Solver2d &S, double D, int nspec, int nw, hipStream_t s, const Ctl *ctl,
                     int skip_mode) {
    if (S.wave) return wlaunch_cols(S, D, nspec, nullptr, s, ctl, skip_mode, nw);
    const int nb = (S.nh + S.cols_per_block - 1) / S.cols_per_block;
    hipLaunchKernelGGL(k_cols_solve, dim3(nb, nspec * nw), dim3(kSolverThreads), cols_lds(S), s, S.spec[0],
                       S.spec[1], S.n1, S.nh, S.ax1.dev, (const double *)S.d_lam1, (const double *)S.d_lam2, D,
                       S.cols_per_block, ctl, skip_mode, 2, nspec, (size_t)S.n1 * S.nh);
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

int solve_cols(const Solver2d &S, double D, int nspec, hipStream_t s, const Ctl *ctl, int skip_mode) {
    if (S.wave) return wlaunch_cols(S, D, nspec, nullptr, s, ctl, skip_mode);
    return cols_launch(S, D, nspec, s, ctl, skip_mode, 2);
}

int solve_cols_forward_only(const Solver2d &S, hipStream_t s) { return cols_launch(S, 0.0, 1, s, nullptr, 0, 0); }

int solve_cols_pass(const Solver2d &S, int nspec, int mode, hipStream_t s) {
    return cols_launch(S, 0.0, nspec, s, nullptr, 0, mode);
}

int solve_cols_inverse_only(const Solver2d &S, hipStream_t s) { return cols_launch(S, 0.0, 1, s, nullptr, 0, 1); }

int fourier_rows_inv(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s) {
    SinkRaw f{dst, S.n2};
    hipLaunchKernelGGL(k_rows_inv<SinkRaw>, dim3(rows_blocks(S)), dim3(kSolverThreads), rows_lds(S), s, spec,
                       spec, f, S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)nullptr, 0, 1,
                       (size_t)0);
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

int solve_rows_inv_unscaled(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s) {
    if (S.wave) return wlaunch_rows_inv(S, spec, spec, SinkRaw{dst, S.n2}, 1, s, (const Ctl *)nullptr, 0);
    return fourier_rows_inv(S, spec, dst, s);
}

int solve_rows_inv_plain_ctl(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s,
                             const Ctl *ctl) {
    SinkScaled f{dst, S.n1, S.n2};
    if (S.wave) return wlaunch_rows_inv(S, spec, spec, f, 1, s, ctl, ctl ? 1 : 0);
    hipLaunchKernelGGL(k_rows_inv<SinkScaled>, dim3(rows_blocks(S)), dim3(kSolverThreads), rows_lds(S), s, spec,
                       spec, f, S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, ctl, ctl ? 1 : 0, 1, (size_t)0);
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

int solve_rows_inv_plain(const Solver2d &S, const double2 *spec, double *dst, hipStream_t s) {
    return solve_rows_inv_plain_ctl(S, spec, dst, s, nullptr);
}

int solve_rows_inv_bcheck(const Solver2d &S, const double2 *spec, double *dst, double thr, Ctl *ctl, int slice,
                          hipStream_t s) {
    SinkBCheck f{dst, S.n1, S.n2, thr, ctl, slice};
    if (S.wave) return wlaunch_rows_inv(S, spec, spec, f, 1, s, (const Ctl *)ctl, 2);
    hipLaunchKernelGGL(k_rows_inv<SinkBCheck>, dim3(rows_blocks(S)), dim3(kSolverThreads), rows_lds(S), s, spec,
                       spec, f, S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)ctl, 2, 1, (size_t)0);
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

}  // namespace qp
