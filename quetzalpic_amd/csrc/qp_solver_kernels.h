// qp_solver_kernels.h -- the three launches of the periodic 2D solve.
// Sources (what the row-forward pass reads) and sinks (what the row-inverse
// pass does with each real output) are small functors so the fused slice
// step can assemble sources and finalise fields inside the transform launches.
//
// Batches: blockIdx.y = wake * nspec + which.  `which` selects spec0 / spec1
// (two independent solves of one wake, e.g. B_y and B_z), `wake` offsets the
// spectra by sstride and the control block by one Ctl; sources and sinks get
// (wake, which) and offset their own grids.
#pragma once

#include "qp_fft_lds.h"

namespace qp {

constexpr int kSolverThreads = 256;

// skip_mode: 0 never, 1 skip when stop_acc (and latch stop_now), 2 skip when stop_now.
// ctl: the control block of this block's wake (or null).
__device__ __forceinline__ bool solver_skip(const Ctl *ctl, int skip_mode) {
    if (!ctl || skip_mode == 0) return false;
    if (skip_mode == 1) {
        if (ctl_skip(ctl)) {
            ctl_latch(const_cast<Ctl *>(ctl));
            return true;
        }
        return false;
    }
    return ctl_skip_b(ctl);
}

struct SrcPlain {
    const double *a;
    int n2;
    __device__ double operator()(int, int, int j, int k) const { return a[(size_t)j * n2 + k]; }
};

// Real output v of the unnormalised c2r; the reference divides by n1 then n2
// (src/System_3d.cpp:361).
struct SinkScaled {
    double *dst;
    int n1, n2;
    __device__ void begin(int) {}
    __device__ void operator()(int, int, int j, int k, double v) const { dst[(size_t)j * n2 + k] = v / n1 / n2; }
};

// Raw unnormalised c2r output (Fourier2d::backward_transform semantics).
struct SinkRaw {
    double *dst;
    int n2;
    __device__ void begin(int) {}
    __device__ void operator()(int, int, int j, int k, double v) const { dst[(size_t)j * n2 + k] = v; }
};

// B store with the stop rule (src/System_3d.cpp:541-552, 569-580).
struct SinkBCheck {
    double *dst;
    int n1, n2;
    double thr;
    Ctl *ctl;
    int slice;
    __device__ void begin(int) {}
    __device__ void operator()(int wk, int, int j, int k, double v) const {
        const double b = v / n1 / n2;
        dst[(size_t)j * n2 + k] = b;
        if (fabs(b) > thr || isnan(b)) {
            ctl[wk].stop_acc = 1;
            ctl[wk].stop_slice = slice;
        }
    }
};

// Rows forward: two real rows (2p, 2p+1) packed as z = x_a + i x_b, complex
// FFT along n2, then X_a[k] = (Z[k] + conj Z[-k])/2, X_b[k] = (Z[k] - conj Z[-k])/(2i).
template <class Src>
__global__ __launch_bounds__(kSolverThreads) void k_rows_fwd(Src src, double2 *spec0, double2 *spec1, int n1,
                                                             int n2, int nh, FftAxisDev ax, int B, const Ctl *ctl,
                                                             int skip_mode, int nspec, size_t sstride) {
    const int which = blockIdx.y % nspec, wk = blockIdx.y / nspec;
    if (solver_skip(ctl ? ctl + wk : nullptr, skip_mode)) return;
    double2 *spec = (which == 0 ? spec0 : spec1) + wk * sstride;
    extern __shared__ double2 lds[];
    double2 *a = lds, *b = lds + (size_t)B * n2, *twl = lds + (size_t)2 * B * n2;
    const int pair0 = blockIdx.x * B;
    lds_load_tw(twl, ax);
    for (int t = threadIdx.x; t < B * n2; t += blockDim.x) {
        const int bb = t / n2, k = t - bb * n2;
        const int ra = 2 * (pair0 + bb), rb = ra + 1;
        double xa = 0.0, xb = 0.0;
        if (ra < n1) xa = src(wk, which, ra, k);
        if (rb < n1) xb = src(wk, which, rb, k);
        a[t] = make_double2(xa, xb);
    }
    __syncthreads();
    const double2 *r = lds_fft(a, b, B, ax, twl, -1);
    for (int t = threadIdx.x; t < B * nh; t += blockDim.x) {
        const int bb = t / nh, k = t - bb * nh;
        const int ra = 2 * (pair0 + bb), rb = ra + 1;
        if (ra >= n1) continue;
        const double2 Z = r[bb * n2 + k];
        const double2 Zm = r[bb * n2 + (n2 - k) % n2];
        const double cr = Zm.x, ci = -Zm.y;
        spec[(size_t)ra * nh + k] = make_double2(0.5 * (Z.x + cr), 0.5 * (Z.y + ci));
        if (rb < n1) {
            const double dr = Z.x - cr, di = Z.y - ci;
            spec[(size_t)rb * nh + k] = make_double2(0.5 * di, -0.5 * dr);
        }
    }
}

// Columns: forward along n1, drop (0,0), divide by the eigenvalue, inverse
// (mode 2); mode 0 = forward only, mode 1 = inverse only (Fourier2d's separate
// forward/backward transforms).  blockIdx.y selects spec0 / spec1 (paired
// solves with the same D).
[[maybe_unused]] static __global__ __launch_bounds__(kSolverThreads) void k_cols_solve(double2 *spec0, double2 *spec1, int n1, int nh,
                                                               FftAxisDev ax, const double *lam1,
                                                               const double *lam2, double D, int C,
                                                               const Ctl *ctl, int skip_mode, int mode, int nspec,
                                                               size_t sstride) {
    const int which = blockIdx.y % nspec, wk = blockIdx.y / nspec;
    if (solver_skip(ctl ? ctl + wk : nullptr, skip_mode)) return;
    extern __shared__ double2 lds[];
    double2 *spec = (which == 0 ? spec0 : spec1) + wk * sstride;
    double2 *a = lds, *b = lds + (size_t)C * n1, *twl = lds + (size_t)2 * C * n1;
    double *lam1l = reinterpret_cast<double *>(twl + n1);
    const int l0 = blockIdx.x * C;
    lds_load_tw(twl, ax);
    for (int t = threadIdx.x; t < n1; t += blockDim.x) lam1l[t] = lam1[t];
    for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
        const int c = t % C, j = t / C, l = l0 + c;
        a[c * n1 + j] = l < nh ? spec[(size_t)j * nh + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (mode == 1) {
        double2 *o = lds_fft(a, b, C, ax, twl, +1);
        for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
            const int c = t % C, j = t / C, l = l0 + c;
            if (l < nh) spec[(size_t)j * nh + l] = o[c * n1 + j];
        }
        return;
    }
    double2 *r = lds_fft(a, b, C, ax, twl, -1);
    if (mode == 0) {
        for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
            const int c = t % C, j = t / C, l = l0 + c;
            if (l < nh) spec[(size_t)j * nh + l] = r[c * n1 + j];
        }
        return;
    }
    for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
        const int c = t / n1, m = t - c * n1, l = l0 + c;
        if (l >= nh) continue;
        double2 v = r[t];
        if (m == 0 && l == 0) {
            v = make_double2(0.0, 0.0);
        } else {
            const double mult = -D + lam1l[m] + lam2[l];
            v.x /= mult;
            v.y /= mult;
        }
        r[t] = v;
    }
    __syncthreads();
    double2 *o = lds_fft(r, r == a ? b : a, C, ax, twl, +1);
    for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
        const int c = t % C, j = t / C, l = l0 + c;
        if (l < nh) spec[(size_t)j * nh + l] = o[c * n1 + j];
    }
}

// Rows inverse (c2r): Z[k] = X[k] + i Y[k] with the Hermitian extension
// X[-k] = conj X[k]; imaginary parts of the k = 0 and k = n2/2 terms ignored
// (FFTW c2r semantics).  Re z -> row 2p, Im z -> row 2p+1.
template <class Sink>
__global__ __launch_bounds__(kSolverThreads) void k_rows_inv(const double2 *spec0, const double2 *spec1, Sink sink,
                                                             int n1, int n2, int nh, FftAxisDev ax, int B,
                                                             const Ctl *ctl, int skip_mode, int nspec,
                                                             size_t sstride) {
    const int which = blockIdx.y % nspec, wk = blockIdx.y / nspec;
    if (solver_skip(ctl ? ctl + wk : nullptr, skip_mode)) return;
    const double2 *spec = (which == 0 ? spec0 : spec1) + wk * sstride;
    sink.begin(wk);
    extern __shared__ double2 lds[];
    double2 *a = lds, *b = lds + (size_t)B * n2, *twl = lds + (size_t)2 * B * n2;
    const int pair0 = blockIdx.x * B;
    lds_load_tw(twl, ax);
    for (int t = threadIdx.x; t < B * n2; t += blockDim.x) {
        const int bb = t / n2, k = t - bb * n2;
        const int ra = 2 * (pair0 + bb), rb = ra + 1;
        double2 z = make_double2(0.0, 0.0);
        if (ra < n1) {
            const bool cj = k >= nh;
            const int kk = cj ? n2 - k : k;
            double2 X = spec[(size_t)ra * nh + kk];
            double2 Y = rb < n1 ? spec[(size_t)rb * nh + kk] : make_double2(0.0, 0.0);
            if (cj) {
                X.y = -X.y;
                Y.y = -Y.y;
            }
            if (kk == 0 || 2 * kk == n2) {
                X.y = 0.0;
                Y.y = 0.0;
            }
            z = make_double2(X.x - Y.y, X.y + Y.x);
        }
        a[t] = z;
    }
    __syncthreads();
    const double2 *r = lds_fft(a, b, B, ax, twl, +1);
    for (int t = threadIdx.x; t < B * n2; t += blockDim.x) {
        const int bb = t / n2, k = t - bb * n2;
        const int ra = 2 * (pair0 + bb), rb = ra + 1;
        if (ra < n1) sink(wk, which, ra, k, r[t].x);
        if (rb < n1) sink(wk, which, rb, k, r[t].y);
    }
}

}  // namespace qp
