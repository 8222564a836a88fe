// qp_fft_lds.h -- batched 1D complex FFT held in LDS (device code).
//
// Stockham autosort, mixed radix, natural order in and out.  Stage with radix
// R after sub-transforms of size Ns:
//   for j in [0, n/R), k = j mod Ns:
//     v_r = in[j + r n/R] * w^{k r (n/(Ns R))}         (w = e^{sign 2 pi i / n})
//     out[(j/Ns) Ns R + k + q Ns] = sum_r v_r w_R^{r q}
// Radix 2 and 4 use closed butterflies; any other radix evaluates each
// output as sum_r in[j + r n/R] * w^{r (k + q Ns) n/(Ns R)} (one twiddle per term).
#pragma once

#include "qp_internal.h"

namespace qp {

__device__ __forceinline__ double2 tw_at(const double2 *tw, int t, int sign) {
    double2 w = tw[t];
    if (sign > 0) w.y = -w.y;
    return w;
}

// Copy the axis twiddle table to LDS (call before the barrier that precedes
// the first lds_fft).
__device__ __forceinline__ void lds_load_tw(double2 *tw_lds, const FftAxisDev &ax) {
    for (int t = threadIdx.x; t < ax.n; t += blockDim.x) tw_lds[t] = ax.tw[t];
}

// Transform `batch` sequences of length ax.n stored contiguously in `a`;
// `b` is scratch of the same size; `twl` is the LDS copy of the twiddles.
// All threads of the block must call it.  Returns the buffer holding the
// result (a or b).
__device__ inline double2 *lds_fft(double2 *a, double2 *b, int batch, const FftAxisDev &ax, const double2 *twl,
                                   int sign) {
    const int n = ax.n;
    int Ns = 1;
    for (int s = 0; s < ax.nstages; ++s) {
        const int R = ax.radix[s];
        const int nR = n / R;
        const int tws = n / (Ns * R);
        if (R == 4) {
            const int total = batch * nR;
            for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
                const int bb = idx / nR, j = idx - bb * nR;
                const double2 *in = a + bb * n;
                double2 *out = b + bb * n;
                const int k = j % Ns;
                double2 v0 = in[j], v1 = in[j + nR], v2 = in[j + 2 * nR], v3 = in[j + 3 * nR];
                if (k != 0) {
                    v1 = cmul(v1, tw_at(twl, k * tws, sign));
                    v2 = cmul(v2, tw_at(twl, 2 * k * tws, sign));
                    v3 = cmul(v3, tw_at(twl, 3 * k * tws, sign));
                }
                const double2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
                const double2 t2 = cadd(v1, v3), t3 = csub(v1, v3);
                // (sign * i) * t3
                const double2 u3 = make_double2(-sign * t3.y, sign * t3.x);
                const int o = (j / Ns) * Ns * 4 + k;
                out[o] = cadd(t0, t2);
                out[o + Ns] = cadd(t1, u3);
                out[o + 2 * Ns] = csub(t0, t2);
                out[o + 3 * Ns] = csub(t1, u3);
            }
        } else if (R == 2) {
            const int total = batch * nR;
            for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
                const int bb = idx / nR, j = idx - bb * nR;
                const double2 *in = a + bb * n;
                double2 *out = b + bb * n;
                const int k = j % Ns;
                double2 v0 = in[j], v1 = in[j + nR];
                if (k != 0) v1 = cmul(v1, tw_at(twl, k * tws, sign));
                const int o = (j / Ns) * Ns * 2 + k;
                out[o] = cadd(v0, v1);
                out[o + Ns] = csub(v0, v1);
            }
        } else {
            const int total = batch * n;
            for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
                const int q = idx % R;
                const int t = idx / R;
                const int bb = t / nR, j = t - bb * nR;
                const double2 *in = a + bb * n;
                double2 *out = b + bb * n;
                const int k = j % Ns;
                const int kq = (k + q * Ns) * tws;
                double2 acc = in[j];
                for (int r = 1; r < R; ++r) {
                    const int ti = (int)(((long long)r * kq) % n);
                    acc = cadd(acc, cmul(in[j + r * nR], tw_at(twl, ti, sign)));
                }
                out[(j / Ns) * Ns * R + k + q * Ns] = acc;
            }
        }
        __syncthreads();
        double2 *tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}

}  // namespace qp
