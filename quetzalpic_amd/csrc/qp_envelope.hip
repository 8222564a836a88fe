// qp_envelope.hip -- laser-envelope advance (no counterpart in the reference).
//
// Equation (x toward the front, slice i at x = i dx, tau the time of the
// moving window, a the complex envelope of the vector potential
// Re[a exp(i k0 (x - t))]):
//     lap_perp a + 2 d_tau (d_x + i k0) a = chi a.
// One step tau -> tau + dt, slices swept from the front (i = nx - 1) down to
// 0 like the wake, with d_i = a+_i - a_i (Crank-Nicolson in tau, d_x as the
// difference toward the front, d_{nx} = 0):
//     (2/dx)(d_{i+1} - d_i) + 2 i k0 d_i = (dt/2)(chi_i - lap_perp)(a+_i + a_i).
// With lap_perp the periodic 5-point Laplacian (eigenvalues lambda =
// (2cos(2 pi j/ny)-2)/dy^2 + (2cos(2 pi k/nz)-2)/dz^2) and c_i the midrange
// (min + max)/2 of chi_i, each transverse mode of slice i reads
//     d_i = M_i(lambda) * [ -(2/dx) d_{i+1} + dt chi_i a_i - dt lap_perp a_i
//                           + (dt/2)(chi_i - c_i) d_i ],
//     M_i = 1 / (alpha + (lambda - c_i) dt/2),  alpha = -2/dx + 2 i k0.
// (chi_i - c_i) d_i couples the modes and is iterated: the first transform
// takes it as 0, each of m further transforms the previous d_i.  For chi >= 0
// the iteration contracts by at most (max - min)/(max + min + 8/(dx dt)) < 1
// for any dt, and a uniform chi is exact in one transform.  The bracket is
// assembled in real space in the row pass (the stencil gives lap_perp a_i
// exactly as the spectrum would), its real and imaginary parts take two real
// transforms, the column pass applies the complex multiplier (mixing the two
// spectra), and the last row pass stores d_i (and, after the last transform
// of the slice, adds it to a_i).  |alpha + (lambda - c_i) dt/2| > 2/dx for
// every mode, so the sweep over the slices is stable for any dt.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "qp_envelope.h"
#include "qp_wave_solver.h"

namespace qp {

namespace {

constexpr int kET = 256;

inline int eblk(size_t n) { return (int)std::min<size_t>((n + kET - 1) / kET, 65535 * 4); }

// a = a0 exp(-4 ((x-x0)^2/xs^2 + (y-y0)^2/ys^2 + (z-z0)^2/zs^2)), real, so
// <a^2> = |a|^2/2 is the Gaussian <a^2> of the wake's laser with the same
// arguments (0.5 a0^2 exp(-8 (...)), src/Config_reader.cpp:114-128).
__global__ void k_env_gauss(double *a, double *asq, size_t vol, size_t nn, int nz, double dx, double dy, double dz,
                            double a0, double ixs2, double iys2, double izs2, double x0, double y0, double z0) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < vol; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t / nn, r = t - i * nn;
        const int j = (int)(r / nz), k = (int)(r - (size_t)j * nz);
        const double x = i * dx - x0, y = j * dy - y0, z = k * dz - z0;
        const double v = a0 * exp(-4.0 * (x * x * ixs2 + y * y * iys2 + z * z * izs2));
        a[t] = v;
        a[vol + t] = 0.0;
        asq[t] = 0.5 * v * v;
    }
}

__global__ void k_env_asq(const double *a, double *asq, size_t vol) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < vol; t += (size_t)gridDim.x * blockDim.x) {
        const double re = a[t], im = a[vol + t];
        asq[t] = 0.5 * (re * re + im * im);
    }
}

// Midrange (min + max)/2 of chi for every slice: one block per slice.
__global__ __launch_bounds__(kET) void k_env_cmid(const double *chi3, double *cm, size_t nn) {
    const double *c = chi3 + (size_t)blockIdx.x * nn;
    double lo = c[0], hi = c[0];
    for (size_t t = threadIdx.x; t < nn; t += kET) {
        lo = fmin(lo, c[t]);
        hi = fmax(hi, c[t]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off));
        hi = fmax(hi, __shfl_xor(hi, off));
    }
    __shared__ double plo[kET / 64], phi[kET / 64];
    if ((threadIdx.x & 63) == 0) {
        plo[threadIdx.x >> 6] = lo;
        phi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kET / 64; ++i) {
            lo = fmin(lo, plo[i]);
            hi = fmax(hi, phi[i]);
        }
        lo = fmin(lo, plo[0]);
        hi = fmax(hi, phi[0]);
        cm[blockIdx.x] = 0.5 * (lo + hi);
    }
}

// Row-pass source of slice e: part `which` (0 re, 1 im) of
// -(2/dx) d_{e+1} + dt (chi a - lap_perp a) + (dt/2)(chi - c_e) d_e.
struct SrcEnv {
    const double *a;      // real parts of all slices (imaginary parts at + vol)
    const double *chi;    // chi of all slices
    const double *inc;    // 2 slots x (re, im) x nn increments
    const double *cmid;   // c of all slices
    size_t vol, nn;
    int nx, n1, n2, use_dcur;
    double idy2, idz2, dt, c2dx;
    int e;
    __device__ double operator()(int, int which, int j, int k) const {
        const double *ac = a + (size_t)which * vol + (size_t)e * nn;
        const int jm = j == 0 ? n1 - 1 : j - 1, jp = j + 1 == n1 ? 0 : j + 1;
        const int km = k == 0 ? n2 - 1 : k - 1, kp = k + 1 == n2 ? 0 : k + 1;
        const size_t t = (size_t)j * n2 + k;
        const double c = ac[t];
        const double lap = (ac[(size_t)jp * n2 + k] - 2.0 * c + ac[(size_t)jm * n2 + k]) * idy2 +
                           (ac[(size_t)j * n2 + kp] - 2.0 * c + ac[(size_t)j * n2 + km]) * idz2;
        const double x = chi[(size_t)e * nn + t];
        double v = dt * (x * c - lap);
        if (e != nx - 1) v -= c2dx * inc[(size_t)((e + 1) & 1) * 2 * nn + (size_t)which * nn + t];
        if (use_dcur) v += 0.5 * dt * (x - cmid[e]) * inc[(size_t)(e & 1) * 2 * nn + (size_t)which * nn + t];
        return v;
    }
};

// Last row pass: d_e (unnormalised c2r output / (n1 n2)) to the increment
// buffer, and onto a_e after the last transform of the slice.
struct SinkEnv {
    double *a;     // all slices (imaginary parts at + vol); null: store d_e only
    double *inc;
    size_t vol, nn;
    int n1, n2, e;
    __device__ void begin(int) {}
    __device__ void operator()(int, int which, int j, int k, double v) const {
        const size_t t = (size_t)j * n2 + k;
        const double d = v / n1 / n2;
        inc[(size_t)(e & 1) * 2 * nn + (size_t)which * nn + t] = d;
        if (a) a[(size_t)which * vol + (size_t)e * nn + t] += d;
    }
};

// M = 1/(alpha + lam dt/2) (lam = lambda - c_i) applied to the spectra X (of
// the real parts) and Y (of the imaginary parts): X' = Re M X - Im M Y,
// Y' = Im M X + Re M Y.
__device__ __forceinline__ void env_mix(double2 &X, double2 &Y, double lam, double ar, double ai, double hdt) {
    const double qr = ar + hdt * lam, qi = ai;
    const double den = qr * qr + qi * qi;
    const double mr = qr / den, mi = -qi / den;
    const double2 x = X, y = Y;
    X = make_double2(mr * x.x - mi * y.x, mr * x.y - mi * y.y);
    Y = make_double2(mi * x.x + mr * y.x, mi * x.y + mr * y.y);
}

// Column pass on the register-resident FFTs: one block holds column l of
// both spectra (column-major spec[l * n1 + j]), forward, mix, inverse.
template <int N1, int E, int TB>
__global__ __launch_bounds__(TB) void kw_cols_env(double2 *spec0, double2 *spec1, int nh, const double2 *tw1,
                                                  const double *lam1, const double *lam2, double ar, double ai,
                                                  double hdt, const double *cmid) {
    using W = WfShape<N1, E, TB>;
    __shared__ double2 sb0[W::LDS], sb1[W::LDS], twl[WfTw<N1, E, TB>::NL];
    const int tid = threadIdx.x;
    const int tt = tid / W::L, j = tid % W::L;
    WfTw<N1, E, TB> tw;
    tw.init(tw1, twl, j);
    const int l = blockIdx.x * W::T + tt;
    const bool valid = l < nh;
    double2 x[E], y[E];
    double lm[E];
    const double2 *c0 = spec0 + (size_t)l * N1, *c1 = spec1 + (size_t)l * N1;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        x[r] = valid ? c0[j + W::L * r] : make_double2(0.0, 0.0);
        y[r] = valid ? c1[j + W::L * r] : make_double2(0.0, 0.0);
        lm[r] = lam1[j + W::L * r];
    }
    const double l2 = (valid ? lam2[l] : 0.0) - *cmid;
    double2 *b0 = sb0 + tt * W::PADN, *b1 = sb1 + tt * W::PADN;
    wave_fft<N1, E, -1>(x, b0, b1, tw, j);
    wave_fft<N1, E, -1>(y, b0, b1, tw, j);
#pragma unroll
    for (int m = 0; m < E; ++m) env_mix(x[m], y[m], lm[m] + l2, ar, ai, hdt);
    wave_fft<N1, E, +1>(x, b0, b1, tw, j);
    wave_fft<N1, E, +1>(y, b0, b1, tw, j);
    if (!valid) return;
    double2 *o0 = spec0 + (size_t)l * N1, *o1 = spec1 + (size_t)l * N1;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        o0[j + W::L * m] = x[m];
        o1[j + W::L * m] = y[m];
    }
}

// Generic path: the multiplier between the forward-only and inverse-only
// column passes, on the row-major spectra spec[j * nh + l].
__global__ void k_env_mix(double2 *spec0, double2 *spec1, int n1, int nh, const double *lam1, const double *lam2,
                          double ar, double ai, double hdt, const double *cmid) {
    const size_t n = (size_t)n1 * nh;
    const double c = *cmid;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / nh), l = (int)(t - (size_t)j * nh);
        double2 X = spec0[t], Y = spec1[t];
        env_mix(X, Y, lam1[j] + lam2[l] - c, ar, ai, hdt);
        spec0[t] = X;
        spec1[t] = Y;
    }
}

#define ELAUNCH(kern, grid, block, lds, ...)                                  \
    do {                                                                      \
        hipLaunchKernelGGL(kern, grid, dim3(block), lds, s, __VA_ARGS__);     \
        QP_HIP_CHECK(hipGetLastError());                                      \
    } while (0)

// One transform per block, as the wake solves' column pass: 256^2 x 1000 slices
// 36.50 -> 35.3 ms per envelope step, 512^2 x 2000 slices 120.5 -> 114.0 ms.
int wlaunch_cols_env(const Solver2d &S, double ar, double ai, double hdt, const double *cmid, hipStream_t s) {
    wf_dispatch(S.n1, [&](auto NC) {
        constexpr int N = decltype(NC)::value, E = WfCfg<N>::E, TB = WfCfg<N>::TBC;
        const dim3 grid(wcols_blocks_t(S.nh, WfShape<N, E, TB>::T));
        hipLaunchKernelGGL((kw_cols_env<N, E, TB>), grid, dim3(TB), 0, s, S.spec[0], S.spec[1], S.nh,
                           (const double2 *)S.ax1.dev.tw, (const double *)S.d_lam1, (const double *)S.d_lam2, ar,
                           ai, hdt, cmid);
    });
    QP_HIP_CHECK(hipGetLastError());
    return QP_OK;
}

// One transform of slice i: d_i from the bracket with the previous iterate
// (use_dcur) or without it, onto a_i when `last`.
int env_transform(Envelope *e, int i, double dt, bool use_dcur, bool last, hipStream_t s) {
    const Solver2d &S = e->S;
    const size_t nn = e->nn(), vol = e->vol();
    const double ar = -2.0 / e->dx, ai = 2.0 * e->k0, hdt = 0.5 * dt;
    SrcEnv src{e->a, e->chi3, e->inc, e->cmid, vol, nn, e->nx, S.n1, S.n2, use_dcur ? 1 : 0,
               1.0 / (e->dy * e->dy), 1.0 / (e->dz * e->dz), dt, 2.0 / e->dx, i};
    SinkEnv snk{last ? e->a : nullptr, e->inc, vol, nn, S.n1, S.n2, i};
    const double *cmid = e->cmid + i;
    if (S.wave) {
        QP_TRY(wlaunch_rows_fwd(S, src, S.spec[0], S.spec[1], 2, s, (const Ctl *)nullptr, 0));
        QP_TRY(wlaunch_cols_env(S, ar, ai, hdt, cmid, s));
        QP_TRY(wlaunch_rows_inv(S, (const double2 *)S.spec[0], (const double2 *)S.spec[1], snk, 2, s,
                                (const Ctl *)nullptr, 0));
        return QP_OK;
    }
    ELAUNCH(k_rows_fwd<SrcEnv>, dim3(rows_blocks(S), 2), kSolverThreads, rows_lds(S), src, S.spec[0], S.spec[1],
            S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)nullptr, 0, 2, (size_t)0);
    QP_TRY(solve_cols_pass(S, 2, 0, s));
    ELAUNCH(k_env_mix, dim3(eblk((size_t)S.n1 * S.nh)), kET, 0, S.spec[0], S.spec[1], S.n1, S.nh,
            (const double *)S.d_lam1, (const double *)S.d_lam2, ar, ai, hdt, cmid);
    QP_TRY(solve_cols_pass(S, 2, 1, s));
    ELAUNCH(k_rows_inv<SinkEnv>, dim3(rows_blocks(S), 2), kSolverThreads, rows_lds(S), (const double2 *)S.spec[0],
            (const double2 *)S.spec[1], snk, S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)nullptr, 0,
            2, (size_t)0);
    return QP_OK;
}

// One step of the whole envelope (every slice, then <a^2>); m corrections of
// the (chi - c) d term per slice.  c_i are set before the first step.
int env_sweep(Envelope *e, double dt, int m, hipStream_t s) {
    const size_t vol = e->vol();
    for (int i = e->nx - 1; i >= 0; --i)
        for (int it = 0; it <= m; ++it) QP_TRY(env_transform(e, i, dt, it > 0, it == m, s));
    ELAUNCH(k_env_asq, dim3(eblk(vol)), kET, 0, (const double *)e->a, e->asq3, vol);
    return QP_OK;
}

void drop_graph(Envelope *e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
}

}  // namespace

// ------------------------------------------------------- wake-side step
// One step of the whole envelope after a wake pass, enqueued on s: midranges
// of chi, then the sweep, replayed from the captured graph of this (dt, m)
// when use_graph, else launched directly.
int envelope_step(Envelope *e, double dt, int m, int use_graph, hipStream_t s) {
    ELAUNCH(k_env_cmid, dim3(e->nx), kET, 0, (const double *)e->chi3, e->cmid, e->nn());
    if (!use_graph) return env_sweep(e, dt, m, s);
    if (!e->gexec || e->g_dtau != dt || e->g_iter != m || e->g_stream != s) {
        drop_graph(e);
        hipGraph_t gr;
        QP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int st = env_sweep(e, dt, m, s);
        const hipError_t ce = hipStreamEndCapture(s, &gr);
        if (st < 0) {
            if (ce == hipSuccess) (void)hipGraphDestroy(gr);
            return st;
        }
        if (ce != hipSuccess) {
            set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
            return QP_ERR_HIP;
        }
        const hipError_t ie = hipGraphInstantiate(&e->gexec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) {
            e->gexec = nullptr;
            set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            return QP_ERR_HIP;
        }
        e->g_dtau = dt;
        e->g_iter = m;
        e->g_stream = s;
    }
    QP_HIP_CHECK(hipGraphLaunch(e->gexec, s));
    return QP_OK;
}

}  // namespace qp

using namespace qp;

#define EARG_CHECK(cond, msg)        \
    do {                             \
        if (!(cond)) {               \
            set_error(msg);          \
            return QP_ERR_ARG;       \
        }                            \
    } while (0)

extern "C" {

int qp_envelope_create(qp_ctx ctx, int nx, int ny, int nz, double dx, double dy, double dz, double k0,
                       qp_envelope *out) {
    EARG_CHECK(ctx && out, "qp_envelope_create: null argument");
    EARG_CHECK(nx >= 1 && ny >= 2 && nz >= 2 && ny <= kMaxN && nz <= kMaxN,
               "qp_envelope_create: need nx >= 1 and 2 <= ny, nz <= 1024");
    EARG_CHECK(dx > 0 && dy > 0 && dz > 0 && k0 > 0, "qp_envelope_create: steps and k0 must be positive");
    *out = nullptr;
    qp_envelope h = new qp_envelope_s();
    Envelope *e = &h->e;
    e->ctx = ctx;
    e->nx = nx;
    e->ny = ny;
    e->nz = nz;
    e->dx = dx;
    e->dy = dy;
    e->dz = dz;
    e->k0 = k0;
    auto fail = [&](int st) {
        qp_envelope_destroy(h);
        return st;
    };
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(QP_ERR_HIP);
    int st = e->S.init(ny, nz, dy, dz, 1);
    if (st != QP_OK) return fail(st);
    const size_t nn = e->nn(), vol = e->vol();
    if (hipMalloc(&e->a, sizeof(double) * 2 * vol) != hipSuccess ||
        hipMalloc(&e->inc, sizeof(double) * 4 * nn) != hipSuccess ||
        hipMalloc(&e->chi3, sizeof(double) * vol) != hipSuccess ||
        hipMalloc(&e->asq3, sizeof(double) * vol) != hipSuccess ||
        hipMalloc(&e->cmid, sizeof(double) * nx) != hipSuccess) {
        set_error("qp_envelope_create: device allocation failed");
        return fail(QP_ERR_ALLOC);
    }
    hipStream_t s = e->stream();
    const void *bufs[] = {e->a, e->chi3, e->asq3};
    const size_t sizes[] = {2 * vol, vol, vol};
    for (int b = 0; b < 3; ++b)
        if (hipMemsetAsync(const_cast<void *>(bufs[b]), 0, sizeof(double) * sizes[b], s) != hipSuccess)
            return fail(QP_ERR_HIP);
    if (hipStreamSynchronize(s) != hipSuccess) return fail(QP_ERR_HIP);
    *out = h;
    return QP_OK;
}

int qp_envelope_destroy(qp_envelope h) {
    if (!h) return QP_OK;
    Envelope *e = &h->e;
    if (e->attached) {
        set_error("qp_envelope_destroy: a wake still reads this envelope (detach or destroy it first)");
        return QP_ERR_STATE;
    }
    if (e->ctx) (void)hipStreamSynchronize(e->stream());
    drop_graph(e);
    e->S.destroy();
    for (double *p : {e->a, e->inc, e->chi3, e->asq3, e->cmid})
        if (p) (void)hipFree(p);

    delete h;
    return QP_OK;
}

int qp_envelope_set(qp_envelope h, const double *re, const double *im) {
    EARG_CHECK(h && re, "qp_envelope_set: null argument");
    Envelope *e = &h->e;
    hipStream_t s = e->stream();
    const size_t vol = e->vol();
    QP_HIP_CHECK(hipMemcpy(e->a, re, sizeof(double) * vol, hipMemcpyHostToDevice));
    if (im) {
        QP_HIP_CHECK(hipMemcpy(e->a + vol, im, sizeof(double) * vol, hipMemcpyHostToDevice));
    } else {
        QP_HIP_CHECK(hipMemsetAsync(e->a + vol, 0, sizeof(double) * vol, s));
    }
    hipLaunchKernelGGL(k_env_asq, dim3(eblk(vol)), dim3(kET), 0, s, (const double *)e->a, e->asq3, vol);
    QP_HIP_CHECK(hipGetLastError());
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

int qp_envelope_set_gaussian(qp_envelope h, double a0, double xsigma, double ysigma, double zsigma, double x0,
                             double y0, double z0) {
    EARG_CHECK(h, "null envelope");
    EARG_CHECK(xsigma > 0 && ysigma > 0 && zsigma > 0, "qp_envelope_set_gaussian: widths must be positive");
    Envelope *e = &h->e;
    hipStream_t s = e->stream();
    const size_t vol = e->vol();
    hipLaunchKernelGGL(k_env_gauss, dim3(eblk(vol)), dim3(kET), 0, s, e->a, e->asq3, vol, e->nn(), e->nz, e->dx,
                       e->dy, e->dz, a0, 1.0 / (xsigma * xsigma), 1.0 / (ysigma * ysigma), 1.0 / (zsigma * zsigma),
                       x0, y0, z0);
    QP_HIP_CHECK(hipGetLastError());
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

int qp_envelope_get(qp_envelope h, double *re, double *im) {
    EARG_CHECK(h, "null envelope");
    Envelope *e = &h->e;
    QP_HIP_CHECK(hipStreamSynchronize(e->stream()));
    const size_t vol = e->vol();
    if (re) QP_HIP_CHECK(hipMemcpy(re, e->a, sizeof(double) * vol, hipMemcpyDeviceToHost));
    if (im) QP_HIP_CHECK(hipMemcpy(im, e->a + vol, sizeof(double) * vol, hipMemcpyDeviceToHost));
    return QP_OK;
}

int qp_envelope_set_chi(qp_envelope h, const double *chi) {
    EARG_CHECK(h, "null envelope");
    Envelope *e = &h->e;
    const size_t vol = e->vol();
    if (chi) {
        QP_HIP_CHECK(hipMemcpy(e->chi3, chi, sizeof(double) * vol, hipMemcpyHostToDevice));
    } else {
        QP_HIP_CHECK(hipMemsetAsync(e->chi3, 0, sizeof(double) * vol, e->stream()));
        QP_HIP_CHECK(hipStreamSynchronize(e->stream()));
    }
    return QP_OK;
}

int qp_envelope_get_chi(qp_envelope h, double *chi) {
    EARG_CHECK(h && chi, "qp_envelope_get_chi: null argument");
    Envelope *e = &h->e;
    QP_HIP_CHECK(hipStreamSynchronize(e->stream()));
    QP_HIP_CHECK(hipMemcpy(chi, e->chi3, sizeof(double) * e->vol(), hipMemcpyDeviceToHost));
    return QP_OK;
}

int qp_envelope_get_a_sqr(qp_envelope h, double *a_sqr) {
    EARG_CHECK(h && a_sqr, "qp_envelope_get_a_sqr: null argument");
    Envelope *e = &h->e;
    QP_HIP_CHECK(hipStreamSynchronize(e->stream()));
    QP_HIP_CHECK(hipMemcpy(a_sqr, e->asq3, sizeof(double) * e->vol(), hipMemcpyDeviceToHost));
    return QP_OK;
}

int qp_envelope_advance(qp_envelope h, double dtau, int steps, int chi_iterations, int use_graph, double *ms) {
    EARG_CHECK(h, "null envelope");
    EARG_CHECK(dtau > 0 && std::isfinite(dtau), "qp_envelope_advance: dtau must be positive");
    EARG_CHECK(steps >= 0, "qp_envelope_advance: steps must be >= 0");
    EARG_CHECK(chi_iterations >= 0 && chi_iterations <= 16, "qp_envelope_advance: chi_iterations in [0, 16]");
    Envelope *e = &h->e;
    hipStream_t s = e->stream();
    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    if (ms) {
        QP_HIP_CHECK(hipEventCreate(&ev.a));
        QP_HIP_CHECK(hipEventCreate(&ev.b));
        QP_HIP_CHECK(hipEventRecord(ev.a, s));
    }
    if (steps > 0) ELAUNCH(k_env_cmid, dim3(e->nx), kET, 0, (const double *)e->chi3, e->cmid, e->nn());
    if (use_graph && steps > 0) {
        if (!e->gexec || e->g_dtau != dtau || e->g_iter != chi_iterations || e->g_stream != s) {
            drop_graph(e);
            hipGraph_t gr;
            QP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int st = env_sweep(e, dtau, chi_iterations, s);
            const hipError_t ce = hipStreamEndCapture(s, &gr);
            if (st < 0) {
                if (ce == hipSuccess) (void)hipGraphDestroy(gr);
                return st;
            }
            if (ce != hipSuccess) {
                set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
                return QP_ERR_HIP;
            }
            const hipError_t ie = hipGraphInstantiate(&e->gexec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ie != hipSuccess) {
                e->gexec = nullptr;
                set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
                return QP_ERR_HIP;
            }
            e->g_dtau = dtau;
            e->g_iter = chi_iterations;
            e->g_stream = s;
        }
        for (int n = 0; n < steps; ++n) QP_HIP_CHECK(hipGraphLaunch(e->gexec, s));
    } else {
        for (int n = 0; n < steps; ++n) QP_TRY(env_sweep(e, dtau, chi_iterations, s));
    }
    if (ms) {
        QP_HIP_CHECK(hipEventRecord(ev.b, s));
        QP_HIP_CHECK(hipEventSynchronize(ev.b));
        float f = 0.f;
        QP_HIP_CHECK(hipEventElapsedTime(&f, ev.a, ev.b));
        *ms = f;
    }
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

}  // extern "C"
