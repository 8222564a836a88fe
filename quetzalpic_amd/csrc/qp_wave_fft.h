// qp_wave_fft.h -- register-resident complex FFTs of length N = 2^p in
// [16, 1024] or N = 3 * 2^p in [48, 384] (device code).
//
// E elements per lane (E in {4, 8, 16} for 2^p, 6 or 12 for 3 * 2^p), L = N/E lanes
// per transform, a block of TB threads carries T = TB/L transforms.  Lane j of a transform
// holds element j + L*m in register m, on entry and on exit (natural order
// both ways), so a forward transform, a pointwise operation and an inverse
// transform chain without an exchange.  Stockham passes: the first has radix
// E and works on the registers; each later pass with sub-transform size Ns
// and radix R = min(E, N/Ns) has butterflies jj = j + L b (b < E/R) reading
// in[jj + r N/R], twiddled by w^{k r N/(Ns R)} (k = jj mod Ns), writing
// out[(jj/Ns) Ns R + k + q Ns]; the last pass lands in registers.  Passes
// exchange through two LDS buffers (one barrier per pass).  For 3 * 2^p the
// first pass has radix E, 6 (3 x 2) or 12 (3 x 4), and the later radices are 4
// or 2, each dividing E (wf_radix).  Twiddles come from the axis table
// e^{-2 pi i t/N}, t < N, in device memory: with E = 4 each lane loads the ones
// it applies into registers at kernel start, and so for 3 * 2^p up to N = 192;
// at N = 384 the block stages the table in LDS (WfTw).
#pragma once

#include <type_traits>

#include "qp_internal.h"

namespace qp {

constexpr bool wf_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
// Lengths the wave FFT handles: 2^p in [16, 1024], 3 * 2^p in [48, 384].
constexpr bool wf_length(int n) {
    return (n >= 16 && n <= 1024 && wf_pow2(n)) || (n >= 48 && n <= 384 && n % 3 == 0 && wf_pow2(n / 3));
}

template <int N, int E, int TB>
struct WfShape {
    static_assert(wf_length(N), "wave FFT length must be 2^p in [16, 1024] or 3 * 2^p in [48, 384]");
    static_assert(N % 3 == 0 ? (E == 6 || E == 12) : (E == 4 || E == 8 || E == 16),
                  "elements per lane must be 4, 8 or 16 (6 or 12 for lengths 3 * 2^p)");
    static constexpr int L = N / E;                      // lanes per transform
    static_assert(TB % 64 == 0 && TB >= L && TB % L == 0, "block must hold whole transforms");
    static constexpr int T = TB / L;                     // transforms per block
    static constexpr int PADN = N + N / 16;              // padded LDS length of one transform
    static constexpr int LDS = T * PADN;
    // two exchange buffers + the twiddle table within the 160 KiB of LDS a gfx950 workgroup may use
    static_assert((2 * LDS + N) * 16 <= 163840, "transforms per block exceed the LDS budget");
};

// One pad slot every 16 elements keeps 16-byte LDS accesses of consecutive
// lanes on distinct banks.
__device__ __forceinline__ int wf_pad(int i) { return i + (i >> 4); }

// cos / sin of 2 pi m / 16 (cos pi/8, sin pi/8, sqrt(1/2)).
constexpr double kC1 = 0.92387953251128675613;
constexpr double kS1 = 0.38268343236508977173;
constexpr double kH = 0.70710678118654752440;
constexpr double cos16(int m) {
    switch (m & 15) {
        case 0: return 1.0;
        case 1: return kC1;
        case 2: return kH;
        case 3: return kS1;
        case 4: return 0.0;
        case 5: return -kS1;
        case 6: return -kH;
        case 7: return -kC1;
        case 8: return -1.0;
        case 9: return -kC1;
        case 10: return -kH;
        case 11: return -kS1;
        case 12: return 0.0;
        case 13: return kS1;
        case 14: return kH;
        default: return kC1;
    }
}
constexpr double sin16(int m) { return cos16(m - 4); }

// cos / sin of 2 pi m / 12 (sqrt(3)/2, 1/2).
constexpr double kR3 = 0.86602540378443864676;
constexpr double cos12(int m) {
    switch (((m % 12) + 12) % 12) {
        case 0: return 1.0;
        case 1: return kR3;
        case 2: return 0.5;
        case 3: return 0.0;
        case 4: return -0.5;
        case 5: return -kR3;
        case 6: return -1.0;
        case 7: return -kR3;
        case 8: return -0.5;
        case 9: return 0.0;
        case 10: return 0.5;
        default: return kR3;
    }
}
constexpr double sin12(int m) { return cos12(m - 3); }

// (S i) * a
template <int S>
__device__ __forceinline__ double2 mul_si(double2 a) {
    return S > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// a * e^{S 2 pi i M / 16}
template <int M, int S>
__device__ __forceinline__ double2 w16(double2 a) {
    constexpr int m = M & 15;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (m == 4) {
        return mul_si<S>(a);
    } else if constexpr (m == 8) {
        return make_double2(-a.x, -a.y);
    } else if constexpr (m == 12) {
        return mul_si<-S>(a);
    } else {
        constexpr double c = cos16(m);
        constexpr double s = S * sin16(m);
        return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// a * e^{S 2 pi i M / 12}
template <int M, int S>
__device__ __forceinline__ double2 w12(double2 a) {
    constexpr int m = M % 12;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (m == 3) {
        return mul_si<S>(a);
    } else if constexpr (m == 6) {
        return make_double2(-a.x, -a.y);
    } else if constexpr (m == 9) {
        return mul_si<-S>(a);
    } else {
        constexpr double c = cos12(m);
        constexpr double s = S * sin12(m);
        return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// Length-3 DFT with w = e^{S 2 pi i / 3} = -1/2 + S i sqrt(3)/2:
// v1, v2 = v0 - (v1 + v2)/2 +- (sqrt(3)/2) S i (v1 - v2).
template <int S>
__device__ __forceinline__ void dft3(double2 &v0, double2 &v1, double2 &v2) {
    const double2 t1 = cadd(v1, v2);
    const double2 t2 = make_double2(v0.x - 0.5 * t1.x, v0.y - 0.5 * t1.y);
    const double2 d = mul_si<S>(csub(v1, v2));
    v0 = cadd(v0, t1);
    v1 = make_double2(t2.x + kR3 * d.x, t2.y + kR3 * d.y);
    v2 = make_double2(t2.x - kR3 * d.x, t2.y - kR3 * d.y);
}

template <int S>
__device__ __forceinline__ void dft4(double2 &v0, double2 &v1, double2 &v2, double2 &v3) {
    const double2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
    const double2 t2 = cadd(v1, v3), t3 = mul_si<S>(csub(v1, v3));
    v0 = cadd(t0, t2);
    v1 = cadd(t1, t3);
    v2 = csub(t0, t2);
    v3 = csub(t1, t3);
}

// In-register DFT of length R in {1, 2, 3, 4, 6, 8, 12, 16}: u[q] <- sum_r u[r] e^{S 2 pi i r q / R}.
// Radix 8 and 16 split as r = r1 + (R/4 or 4) r2 with internal twiddles; radix
// 12 as r = r1 + 4 r2, q = q2 + 3 q1: length-3 DFTs over r2, twiddles
// e^{S 2 pi i r1 q2 / 12}, length-4 DFTs over r1; radix 6 the same way with
// r = r1 + 2 r2 and length-2 DFTs over r1.
template <int R, int S>
__device__ __forceinline__ void dft(double2 (&u)[R]) {
    if constexpr (R == 3) {
        dft3<S>(u[0], u[1], u[2]);
    } else if constexpr (R == 6) {
        double2 a0 = u[0], a1 = u[2], a2 = u[4], b0 = u[1], b1 = u[3], b2 = u[5];
        dft3<S>(a0, a1, a2);
        dft3<S>(b0, b1, b2);
        b1 = w12<2, S>(b1);  // e^{S 2 pi i q2 / 6} for r1 = 1
        b2 = w12<4, S>(b2);
        u[0] = cadd(a0, b0);
        u[3] = csub(a0, b0);
        u[1] = cadd(a1, b1);
        u[4] = csub(a1, b1);
        u[2] = cadd(a2, b2);
        u[5] = csub(a2, b2);
    } else if constexpr (R == 12) {
        double2 a[4][3];
#pragma unroll
        for (int r1 = 0; r1 < 4; ++r1) {
            a[r1][0] = u[r1];
            a[r1][1] = u[r1 + 4];
            a[r1][2] = u[r1 + 8];
            dft3<S>(a[r1][0], a[r1][1], a[r1][2]);
        }
        a[1][1] = w12<1, S>(a[1][1]);
        a[1][2] = w12<2, S>(a[1][2]);
        a[2][1] = w12<2, S>(a[2][1]);
        a[2][2] = w12<4, S>(a[2][2]);
        a[3][1] = w12<3, S>(a[3][1]);
        a[3][2] = w12<6, S>(a[3][2]);
#pragma unroll
        for (int q2 = 0; q2 < 3; ++q2) {
            double2 c0 = a[0][q2], c1 = a[1][q2], c2 = a[2][q2], c3 = a[3][q2];
            dft4<S>(c0, c1, c2, c3);
            u[q2] = c0;
            u[q2 + 3] = c1;
            u[q2 + 6] = c2;
            u[q2 + 9] = c3;
        }
    } else if constexpr (R == 2) {
        const double2 a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    } else if constexpr (R == 4) {
        dft4<S>(u[0], u[1], u[2], u[3]);
    } else if constexpr (R == 8) {
        // r = r1 + 2 r2, q = q2 + 4 q1
        double2 a0 = u[0], a1 = u[2], a2 = u[4], a3 = u[6];
        double2 b0 = u[1], b1 = u[3], b2 = u[5], b3 = u[7];
        dft4<S>(a0, a1, a2, a3);
        dft4<S>(b0, b1, b2, b3);
        b1 = w16<2, S>(b1);
        b2 = w16<4, S>(b2);
        b3 = w16<6, S>(b3);
        u[0] = cadd(a0, b0);
        u[4] = csub(a0, b0);
        u[1] = cadd(a1, b1);
        u[5] = csub(a1, b1);
        u[2] = cadd(a2, b2);
        u[6] = csub(a2, b2);
        u[3] = cadd(a3, b3);
        u[7] = csub(a3, b3);
    } else if constexpr (R == 16) {
        // r = r1 + 4 r2, q = q2 + 4 q1
        double2 a[4][4];
#pragma unroll
        for (int r1 = 0; r1 < 4; ++r1) {
            a[r1][0] = u[r1];
            a[r1][1] = u[r1 + 4];
            a[r1][2] = u[r1 + 8];
            a[r1][3] = u[r1 + 12];
            dft4<S>(a[r1][0], a[r1][1], a[r1][2], a[r1][3]);
        }
        a[1][1] = w16<1, S>(a[1][1]);
        a[1][2] = w16<2, S>(a[1][2]);
        a[1][3] = w16<3, S>(a[1][3]);
        a[2][1] = w16<2, S>(a[2][1]);
        a[2][2] = w16<4, S>(a[2][2]);
        a[2][3] = w16<6, S>(a[2][3]);
        a[3][1] = w16<3, S>(a[3][1]);
        a[3][2] = w16<6, S>(a[3][2]);
        a[3][3] = w16<9, S>(a[3][3]);
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) {
            double2 c0 = a[0][q2], c1 = a[1][q2], c2 = a[2][q2], c3 = a[3][q2];
            dft4<S>(c0, c1, c2, c3);
            u[q2] = c0;
            u[q2 + 4] = c1;
            u[q2 + 8] = c2;
            u[q2 + 12] = c3;
        }
    }
}


// Orders the LDS exchanges of one transform.  A transform of L <= 64 lanes
// lies in one wavefront (L divides 64), whose LDS instructions execute in
// program order, so a wavefront-scope fence (which keeps the compiler from
// moving LDS accesses across it) is enough; wider transforms need the block
// barrier.
template <int L>
__device__ __forceinline__ void wf_sync() {
    if constexpr (L <= 64) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// Radix of the pass after sub-transforms of size NS: for 2^p lengths the
// remaining length capped at E; for 3 * 2^p (E = 6 or 12) the remaining length
// when it divides E (the last pass), else 4 or 2 (every radix must divide E).
constexpr int wf_radix(int N, int E, int NS) {
    const int rem = N / NS;
    if (E % 3 != 0) return rem < E ? rem : E;
    if (rem <= E && E % rem == 0) return rem;
    return (rem % 4 == 0 && E % 4 == 0) ? 4 : 2;
}

// Twiddles one lane applies in the passes after the first (sub-transforms of
// size NS upward).
constexpr int wf_ntw(int N, int E, int NS) {
    return NS >= N ? 0 : (E / wf_radix(N, E, NS)) * (wf_radix(N, E, NS) - 1) + wf_ntw(N, E, NS * wf_radix(N, E, NS));
}

// Loads, in pass order, the twiddles e^{-2 pi i t/N} lane j applies: slot
// OFF + b (R - 1) + r - 1 of a pass holds the one of its butterfly b, input r.
template <int N, int E, int NS, int OFF, int NT>
__device__ __forceinline__ void wf_tw_load(double2 (&t)[NT], const double2 *__restrict__ g, int j) {
    if constexpr (NS < N) {
        constexpr int L = N / E;
        constexpr int R = wf_radix(N, E, NS);
        constexpr int NB = E / R;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int k = (j + L * b) % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) t[OFF + b * (R - 1) + r - 1] = g[k * r * (N / (NS * R))];
        }
        wf_tw_load<N, E, NS * R, OFF + NB * (R - 1)>(t, g, j);
    }
}

// Twiddle source of wave_fft.  Every lane loads the twiddles it applies into
// registers at kernel start, next to its data loads: no LDS table and no
// barrier to publish it.  That holds 3 per pass with 4 elements per lane, and
// up to 15 in all for lengths 3 * 2^p <= 192 with 6 per lane (18 with 12); for
// 384 the block stages the N-entry table in LDS instead.
template <int N, int E, int TB>
struct WfTw {
    static constexpr bool kRegs = E == 4 || N <= 192;
    static constexpr int NT = kRegs ? wf_ntw(N, E, E) : 1;
    static constexpr int NL = kRegs ? 1 : N;  // entries of the LDS table
    double2 reg[NT > 0 ? NT : 1];
    const double2 *lds;
    // g: the table e^{-2 pi i t/N} in global memory; table: NL entries of LDS.
    // Every thread of the block calls it.
    __device__ __forceinline__ void init(const double2 *__restrict__ g, double2 *table, int j) {
        if constexpr (kRegs) {
            wf_tw_load<N, E, E, 0>(reg, g, j);
        } else {
            for (int t = threadIdx.x; t < N; t += TB) table[t] = g[t];
            __syncthreads();  // twiddle table in LDS for every transform of the block
            lds = table;
        }
    }
    // e^{S 2 pi i t/N}, from register slot s or table entry t.
    template <int S>
    __device__ __forceinline__ double2 get(int s, int t) const {
        double2 w;
        if constexpr (kRegs) {
            w = reg[s];
        } else {
            w = lds[t];
        }
        if constexpr (S > 0) w.y = -w.y;
        return w;
    }
};

// Passes after the first: bin holds sub-transforms of size NS.
template <int N, int E, int S, int NS, int OFF, class TW>
__device__ __forceinline__ void wf_passes(double2 (&v)[E], double2 *bin, double2 *bout, const TW &tw, int j) {
    constexpr int L = N / E;
    constexpr int REM = N / NS;
    constexpr int R = wf_radix(N, E, NS);
    constexpr bool LAST = (R == REM);
    constexpr int NB = E / R;
    static_assert(E % R == 0 && REM % R == 0, "pass radix must divide E and the remaining length");
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int jj = j + L * b;
        const int k = jj % NS;
        double2 u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = bin[wf_pad(jj + r * (N / R))];
        if (k != 0) {
#pragma unroll
            for (int r = 1; r < R; ++r)
                u[r] = cmul(u[r], tw.template get<S>(OFF + b * (R - 1) + r - 1, k * r * (N / (NS * R))));
        }
        dft<R, S>(u);
        if constexpr (LAST) {
#pragma unroll
            for (int q = 0; q < R; ++q) v[b + NB * q] = u[q];
        } else {
            const int base = (jj / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) bout[wf_pad(base + q * NS)] = u[q];
        }
    }
    if constexpr (!LAST) {
        wf_sync<L>();
        wf_passes<N, E, S, NS * R, OFF + NB * (R - 1)>(v, bout, bin, tw, j);
    }
}

// Transform the E registers of every lane (see the header comment).  buf0 /
// buf1: this transform's LDS regions (PADN each); tw: the WfTw<N, E, TB> the
// block initialised.  Every thread of the block must call it.
template <int N, int E, int S, class TW>
__device__ __forceinline__ void wave_fft(double2 (&v)[E], double2 *buf0, double2 *buf1, const TW &tw, int j) {
    constexpr int L = N / E;
    dft<E, S>(v);
    if constexpr (N > E) {
        wf_sync<L>();  // earlier readers of buf0 are done
#pragma unroll
        for (int q = 0; q < E; ++q) buf0[wf_pad(E * j + q)] = v[q];
        wf_sync<L>();
        wf_passes<N, E, S, E, 0>(v, buf0, buf1, tw, j);
    }
}

// Run f(std::integral_constant<int, N>) for the wave-FFT length n; returns
// false when n is not a wave-FFT length (wf_length).
template <class F>
inline bool wf_dispatch(int n, F &&f) {
    switch (n) {
        case 16: f(std::integral_constant<int, 16>{}); return true;
        case 32: f(std::integral_constant<int, 32>{}); return true;
        case 48: f(std::integral_constant<int, 48>{}); return true;
        case 64: f(std::integral_constant<int, 64>{}); return true;
        case 96: f(std::integral_constant<int, 96>{}); return true;
        case 128: f(std::integral_constant<int, 128>{}); return true;
        case 192: f(std::integral_constant<int, 192>{}); return true;
        case 256: f(std::integral_constant<int, 256>{}); return true;
        case 384: f(std::integral_constant<int, 384>{}); return true;
        case 512: f(std::integral_constant<int, 512>{}); return true;
        case 1024: f(std::integral_constant<int, 1024>{}); return true;
        default: return false;
    }
}

inline bool wf_supported(int n) { return wf_length(n); }

}  // namespace qp
