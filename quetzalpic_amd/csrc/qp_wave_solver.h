// qp_wave_solver.h -- the three launches of the periodic 2D solve built on the
// register-resident FFTs of qp_wave_fft.h (both grid sides 2^p in [16, 1024]).
//
// Spectrum layout: column-major half spectrum spec[l * n1 + j] (l < nh =
// n2/2 + 1 along the row axis, j < n1), so the column pass reads and writes
// contiguous columns and the row passes move 2T consecutive rows per
// frequency.  Same sources / sinks and the same batch convention
// (blockIdx.y = wake * nspec + which) as the generic kernels of
// qp_solver_kernels.h.
#pragma once

#include "qp_solver_kernels.h"
#include "qp_wave_fft.h"

namespace qp {

// Launch shape per transform length: 4 elements per lane (more waves in
// flight beat fewer exchanges: at 256^2 the psi solve in a dependent chain
// measured 17.7 us with E = 4 against 23.1 us with E = 8 and 34.2 us with
// E = 16) and the threads per block that measured best at 256^2, 512^2 and
// 1024^2 (128, 256, 512); smaller grids use one wavefront per block.  At 1024
// two transforms per block (69.6 KB of LDS, within gfx950's 160 KB per
// workgroup) make the row passes' column-major spectrum accesses 64 B
// contiguous instead of 32 B (a 1024^2 x 4000-slice beam-driven pass 5.86 ->
// 5.74 s); the same at 512 (four transforms) was slower (a 512^2 x 2000-slice
// pass 665 -> 733 ms).
// Elements per lane of the 3 * 2^p lengths: 6 (a radix-6 first pass, then
// radix 2) gives twice the wavefronts of 12 (radix 12, then 4) at half the work
// per lane; these launches have few blocks, so more, shorter waves win (the
// 64-wake C4 pass in three chains 1555 -> 1480 ms, as one group 1808 -> 1779 ms).
constexpr int kWfE3 = 6;
template <int N>
struct WfCfg {
    static constexpr int E = N % 3 == 0 ? kWfE3 : 4;
    static constexpr int TB = N == 1024 ? 512 : (N == 512 ? 256 : (N == 256 ? 128 : 64));
    // Column pass of the wake solves: one transform per block (columns are
    // contiguous, so the row passes' reason for several does not apply; more
    // blocks, and a smaller border-mean reduction).  A 256^2 x 1000-slice wake
    // 269.5 -> 266.6 ms per pass, 512^2 x 2000 slices 667 -> 662 ms, a
    // 1024^2 x 4000-slice beam-driven pass 5.69 -> 5.57 s.
    static constexpr int TBC = N == 256 ? 64 : (N == 512 ? 128 : (N == 1024 ? 256 : TB));
};

// Sum of three values over a block of TB threads (every thread gets the sums).
template <int TB>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    if constexpr (TB > 64) {
        __shared__ double part[3][TB / 64];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        __syncthreads();
        if (lane == 0) {
            part[0][wv] = a;
            part[1][wv] = b;
            part[2][wv] = c;
        }
        __syncthreads();
        a = b = c = 0.0;
        for (int i = 0; i < TB / 64; ++i) {
            a += part[0][i];
            b += part[1][i];
            c += part[2][i];
        }
    }
}

// Rows forward: rows (2p, 2p+1) packed as z = x_a + i x_b, complex FFT along
// n2, X_a[k] = (Z[k] + conj Z[-k])/2, X_b[k] = (Z[k] - conj Z[-k])/(2i).
// The bodies of the three solver kernels take their block coordinates (bx,
// by) and LDS buffers as arguments: the plain kernels pass blockIdx and their
// own static LDS, the group launch (several batches of different grid sizes
// in one launch) the coordinates inside one batch and a shared dynamic buffer.
template <int N2, int E, int TB, class Src>
__device__ __forceinline__ void wrows_fwd_body(const Src &src, double2 *spec0, double2 *spec1, int n1,
                                               const double2 *tw2, const Ctl *ctl, int skip_mode, int nspec,
                                               size_t sstride, int bx, int by, double2 *sb0, double2 *sb1,
                                               double2 *twl) {
    using W = WfShape<N2, E, TB>;
    const int which = by % nspec, wk = by / nspec;
    if (solver_skip(ctl ? ctl + wk : nullptr, skip_mode)) return;
    double2 *spec = (which == 0 ? spec0 : spec1) + wk * sstride;
    const int tid = threadIdx.x;
    const int tt = tid / W::L, j = tid % W::L;
    WfTw<N2, E, TB> tw;
    tw.init(tw2, twl, j);
    const int ra = 2 * (bx * W::T + tt), rb = ra + 1;
    double2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int k = j + W::L * r;
        double xa = 0.0, xb = 0.0;
        if (ra < n1) xa = src(wk, which, ra, k);
        if (rb < n1) xb = src(wk, which, rb, k);
        v[r] = make_double2(xa, xb);
    }
    wave_fft<N2, E, -1>(v, sb0 + tt * W::PADN, sb1 + tt * W::PADN, tw, j);
    // Z in natural order to LDS (after the barrier no lane reads sb0 any more)
    double2 *zb = sb0;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) zb[tt * W::PADN + wf_pad(j + W::L * m)] = v[m];
    __syncthreads();
    // outputs: thread = t2 + T kk writes frequencies kk + L i of row pair t2
    const int t2 = tid % W::T, kk = tid / W::T;
    const int qa = 2 * (bx * W::T + t2), qb = qa + 1;
    if (qa >= n1) return;
    const double2 *z = zb + t2 * W::PADN;
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) {
        const int k = kk + W::L * i;
        if (k > N2 / 2) break;
        const double2 Z = z[wf_pad(k)];
        const double2 Zm = z[wf_pad(k == 0 ? 0 : N2 - k)];
        const double cr = Zm.x, ci = -Zm.y;
        spec[(size_t)k * n1 + qa] = make_double2(0.5 * (Z.x + cr), 0.5 * (Z.y + ci));
        if (qb < n1) spec[(size_t)k * n1 + qb] = make_double2(0.5 * (Z.y - ci), -0.5 * (Z.x - cr));
    }
}

template <int N2, int E, int TB, class Src>
__global__ __launch_bounds__(TB) void kw_rows_fwd(Src src, double2 *spec0, double2 *spec1, int n1,
                                                  const double2 *tw2, const Ctl *ctl, int skip_mode, int nspec,
                                                  size_t sstride) {
    using W = WfShape<N2, E, TB>;
    __shared__ double2 sb0[W::LDS], sb1[W::LDS], twl[WfTw<N2, E, TB>::NL];
    wrows_fwd_body<N2, E, TB, Src>(src, spec0, spec1, n1, tw2, ctl, skip_mode, nspec, sstride, blockIdx.x,
                                   blockIdx.y, sb0, sb1, twl);
}

// Columns: forward along n1, (0,0) := 0, divide by -D + lam1[m] + lam2[l],
// inverse.  With bpart != nullptr (psi solves) also the border-mean partial
// sums of the later c2r output x (unnormalised):
//   sum_j x[j][0]      = sum_l c_l n1 Re H[0][l]
//   x[0][0]            = sum_l c_l Re G[0][l],  G[0][l] = sum_m H[m][l]
//   sum_k x[0][k]      = n2 Re G[0][0]
// (H = divided spectrum, G = its inverse column transform, c_l = 1 for l = 0
// and l = n2/2, else 2); per block: bpart[2b] = first sum, bpart[2b+1] =
// second, and block 0 writes Re G[0][0] to bpart[2 gridDim.x] (each wake of
// a batch at its own offset bstride).
template <int N1, int E, int TB>
__device__ __forceinline__ void wcols_body(double2 *spec0, double2 *spec1, int nh, int n2, const double2 *tw1,
                                           const double *lam1, const double *lam2, double D, double *bpart,
                                           const Ctl *ctl, int skip_mode, int nspec, size_t sstride,
                                           size_t bstride, int bx, int by, int gx, double2 *sb0, double2 *sb1,
                                           double2 *twl) {
    using W = WfShape<N1, E, TB>;
    const int which = by % nspec, wk = by / nspec;
    if (solver_skip(ctl ? ctl + wk : nullptr, skip_mode)) return;
    double2 *spec = (which == 0 ? spec0 : spec1) + wk * sstride;
    if (bpart) bpart += wk * bstride;
    const int tid = threadIdx.x;
    const int tt = tid / W::L, j = tid % W::L;
    WfTw<N1, E, TB> tw;
    tw.init(tw1, twl, j);
    const int l = bx * W::T + tt;
    const bool valid = l < nh;
    double2 v[E];
    double lm[E];
    const double2 *col = spec + (size_t)l * N1;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        v[r] = valid ? col[j + W::L * r] : make_double2(0.0, 0.0);
        lm[r] = lam1[j + W::L * r];
    }
    const double l2 = valid ? lam2[l] : 0.0;
    double2 *b0 = sb0 + tt * W::PADN, *b1 = sb1 + tt * W::PADN;
    wave_fft<N1, E, -1>(v, b0, b1, tw, j);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        if (!valid || (l == 0 && j == 0 && m == 0)) {
            v[m] = make_double2(0.0, 0.0);
        } else {
            const double inv = 1.0 / (-D + lm[m] + l2);  // one division per complex value
            v[m].x *= inv;
            v[m].y *= inv;
        }
    }
    if (bpart) {
        const double wl = (l == 0 || 2 * l == n2) ? 1.0 : 2.0;
        double g = 0.0;
#pragma unroll
        for (int m = 0; m < E; ++m) g += v[m].x;
        double s1 = (valid && j == 0) ? wl * (double)N1 * v[0].x : 0.0;
        double s2 = valid ? wl * g : 0.0;
        double g00 = (l == 0) ? g : 0.0;
        block_sum3<TB>(s1, s2, g00);
        if (tid == 0) {
            bpart[2 * bx] = s1;
            bpart[2 * bx + 1] = s2;
            if (bx == 0) bpart[2 * gx] = g00;
        }
    }
    wave_fft<N1, E, +1>(v, b0, b1, tw, j);
    if (!valid) return;
    double2 *out = spec + (size_t)l * N1;
#pragma unroll
    for (int m = 0; m < E; ++m) out[j + W::L * m] = v[m];
}

template <int N1, int E, int TB>
__global__ __launch_bounds__(TB) void kw_cols(double2 *spec0, double2 *spec1, int nh, int n2, const double2 *tw1,
                                              const double *lam1, const double *lam2, double D, double *bpart,
                                              const Ctl *ctl, int skip_mode, int nspec, size_t sstride,
                                              size_t bstride) {
    using W = WfShape<N1, E, TB>;
    __shared__ double2 sb0[W::LDS], sb1[W::LDS], twl[WfTw<N1, E, TB>::NL];
    wcols_body<N1, E, TB>(spec0, spec1, nh, n2, tw1, lam1, lam2, D, bpart, ctl, skip_mode, nspec, sstride, bstride,
                          blockIdx.x, blockIdx.y, gridDim.x, sb0, sb1, twl);
}

// Rows inverse (c2r): Z[k] = X[k] + i Y[k] with X[-k] = conj X[k]; the
// imaginary parts of the k = 0 and k = n2/2 terms are ignored (FFTW c2r
// semantics).  Re z -> row 2p, Im z -> row 2p+1, handed to the sink.
template <int N2, int E, int TB, class Sink, bool STAGE>
__device__ __forceinline__ void wrows_inv_body(const double2 *spec0, const double2 *spec1, Sink sink, int n1,
                                               const double2 *tw2, const Ctl *ctl, int skip_mode, int nspec,
                                               size_t sstride, int bx, int by, double2 *sb0, double2 *sb1,
                                               double2 *twl) {
    using W = WfShape<N2, E, TB>;
    const int which = by % nspec, wk = by / nspec;
    if (solver_skip(ctl ? ctl + wk : nullptr, skip_mode)) return;
    const double2 *spec = (which == 0 ? spec0 : spec1) + wk * sstride;
    sink.begin(wk);
    const int tid = threadIdx.x;
    const int tt = tid / W::L, j = tid % W::L;
    WfTw<N2, E, TB> tw;
    tw.init(tw2, twl, j);
    const int ra = 2 * (bx * W::T + tt), rb = ra + 1;
    constexpr int NHF = N2 / 2 + 1;  // frequencies of the half spectrum
    // STAGE (lengths >= 32): the block first copies its 2T rows of the
    // column-major half spectrum to LDS, neighbouring lanes reading
    // neighbouring rows of one frequency (one contiguous run per frequency
    // instead of one 16-byte piece per lane).  Row pair t lands in sb1 + t PADN
    // (row ra at [0, NHF), rb at [NHF, 2 NHF)); its lanes read it into
    // registers before wave_fft's first write to sb1.  Rows beyond n1 are
    // staged as zeros.  The extra barrier and LDS round trip pay off only in
    // large launches (the launcher decides).
    constexpr bool staged = STAGE && 2 * NHF <= W::PADN;
    if constexpr (staged) {
        constexpr int R2 = 2 * W::T;
        const int row0 = 2 * bx * W::T;
        for (int e = tid; e < R2 * NHF; e += TB) {
            const int r2 = e % R2, f = e / R2;
            const int row = row0 + r2;
            sb1[(r2 >> 1) * W::PADN + (r2 & 1) * NHF + f] =
                row < n1 ? spec[(size_t)f * n1 + row] : make_double2(0.0, 0.0);
        }
        __syncthreads();
    }
    const double2 *st = sb1 + tt * W::PADN;
    double2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int k = j + W::L * r;
        const bool cj = k > N2 / 2;
        const int kk = cj ? N2 - k : k;
        double2 X = make_double2(0.0, 0.0), Y = X;
        if constexpr (staged) {
            X = st[kk];
            Y = st[NHF + kk];
        } else {
            if (ra < n1) X = spec[(size_t)kk * n1 + ra];
            if (rb < n1) Y = spec[(size_t)kk * n1 + rb];
        }
        if (cj) {
            X.y = -X.y;
            Y.y = -Y.y;
        }
        if (kk == 0 || 2 * kk == N2) {
            X.y = 0.0;
            Y.y = 0.0;
        }
        v[r] = make_double2(X.x - Y.y, X.y + Y.x);
    }
    wave_fft<N2, E, +1>(v, sb0 + tt * W::PADN, sb1 + tt * W::PADN, tw, j);
    if (ra >= n1) return;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int k = j + W::L * m;
        sink(wk, which, ra, k, v[m].x);
        if (rb < n1) sink(wk, which, rb, k, v[m].y);
    }
}

template <int N2, int E, int TB, class Sink, bool STAGE>
__global__ __launch_bounds__(TB) void kw_rows_inv(const double2 *spec0, const double2 *spec1, Sink sink, int n1,
                                                  const double2 *tw2, const Ctl *ctl, int skip_mode, int nspec,
                                                  size_t sstride) {
    using W = WfShape<N2, E, TB>;
    __shared__ double2 sb0[W::LDS], sb1[W::LDS], twl[WfTw<N2, E, TB>::NL];
    wrows_inv_body<N2, E, TB, Sink, STAGE>(spec0, spec1, sink, n1, tw2, ctl, skip_mode, nspec, sstride, blockIdx.x,
                                           blockIdx.y, sb0, sb1, twl);
}

// ------------------------------------------------------------ launchers
// Blocks along x of the row passes / the column pass for transforms per block T.
inline int wrows_blocks_t(int n1, int T) { return ((n1 + 1) / 2 + T - 1) / T; }
inline int wcols_blocks_t(int nh, int T) { return (nh + T - 1) / T; }

inline int wcols_blocks(int n1, int nh) {
    int nb = 0;
    wf_dispatch(n1, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        nb = wcols_blocks_t(nh, WfShape<N, WfCfg<N>::E, WfCfg<N>::TBC>::T);
    });
    return nb;
}

// nw wakes of one batch: spectra of wake w at spec0/spec1 + w * n1 * nh, control
// blocks at ctl + w.
template <class Src>
int wlaunch_rows_fwd(const Solver2d &S, const Src &src, double2 *spec0, double2 *spec1, int nspec, hipStream_t s,
                     const Ctl *ctl, int skip_mode, int nw = 1) {
    const size_t sstride = (size_t)S.n1 * S.nh;
    wf_dispatch(S.n2, [&](auto NC) {
        constexpr int N = decltype(NC)::value, E = WfCfg<N>::E, TB = WfCfg<N>::TB;
        const dim3 grid(wrows_blocks_t(S.n1, WfShape<N, E, TB>::T), nspec * nw);
        hipLaunchKernelGGL((kw_rows_fwd<N, E, TB, Src>), grid, dim3(TB), 0, s, src, spec0, spec1, S.n1,
                           (const double2 *)S.ax2.dev.tw, ctl, skip_mode, nspec, sstride);
    });
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

inline int wlaunch_cols(const Solver2d &S, double D, int nspec, double *bpart, hipStream_t s, const Ctl *ctl,
                        int skip_mode, int nw = 1, size_t bstride = 0) {
    const size_t sstride = (size_t)S.n1 * S.nh;
    wf_dispatch(S.n1, [&](auto NC) {
        constexpr int N = decltype(NC)::value, E = WfCfg<N>::E, TB = WfCfg<N>::TBC;
        const dim3 grid(wcols_blocks_t(S.nh, WfShape<N, E, TB>::T), nspec * nw);
        hipLaunchKernelGGL((kw_cols<N, E, TB>), grid, dim3(TB), 0, s, S.spec[0], S.spec[1], S.nh, S.n2,
                           (const double2 *)S.ax1.dev.tw, (const double *)S.d_lam1, (const double *)S.d_lam2, D,
                           bpart, ctl, skip_mode, nspec, sstride, bstride);
    });
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

template <class Sink>
int wlaunch_rows_inv(const Solver2d &S, const double2 *spec0, const double2 *spec1, const Sink &sink, int nspec,
                     hipStream_t s, const Ctl *ctl, int skip_mode, int nw = 1) {
    const size_t sstride = (size_t)S.n1 * S.nh;
    wf_dispatch(S.n2, [&](auto NC) {
        constexpr int N = decltype(NC)::value, E = WfCfg<N>::E, TB = WfCfg<N>::TB;
        const dim3 grid(wrows_blocks_t(S.n1, WfShape<N, E, TB>::T), nspec * nw);
        // Staged spectrum reads for large launches: a 1024^2 x 4000-slice
        // beam-driven pass 5.75 -> 5.68 s, a batch of ten 256^2 wakes 648.5 -> 640.6 ms per pass.
        // One 256^2 wake (64 or 128 blocks, latency bound) was slower staged,
        // 268.7 -> 282.2 ms per pass, so small launches read directly.
        if (N >= 1024 || grid.x * grid.y >= 1024)
            hipLaunchKernelGGL((kw_rows_inv<N, E, TB, Sink, true>), grid, dim3(TB), 0, s, spec0, spec1, sink, S.n1,
                               (const double2 *)S.ax2.dev.tw, ctl, skip_mode, nspec, sstride);
        else
            hipLaunchKernelGGL((kw_rows_inv<N, E, TB, Sink, false>), grid, dim3(TB), 0, s, spec0, spec1, sink, S.n1,
                               (const double2 *)S.ax2.dev.tw, ctl, skip_mode, nspec, sstride);
    });
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

}  // namespace qp
