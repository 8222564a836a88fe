// qp_wake_fp64.hip -- the plain FP64 slice loop on the device.
//
// One launch per loop of System_3d::solve_wakefield (src/System_3d.cpp:238-627),
// FP64 structure-of-arrays particles, global FP64 atomics for the CIC
// deposits, the three-launch periodic solve.  Arithmetic follows the
// reference's operation order (this file is built with -ffp-contract=off) so
// the result agrees with the FP64 oracle at rounding level.  It is the
// baseline the fused mixed path is measured against.
#include <cstring>

#include "qp_solver_kernels.h"
#include "qp_wake.h"

namespace qp {

constexpr int kT = 256;
static inline int nblk(size_t n) { return (int)((n + kT - 1) / kT); }

#define SKIP_OR_LATCH(ctl)        \
    if (ctl_skip(ctl)) {          \
        ctl_latch(ctl);           \
        return;                   \
    }

// ------------------------------------------------------------- grid kernels
__global__ void k_copy(double *dst, const double *src, size_t n, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

__global__ void k_zero2(double *a, double *b, size_t n, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) {
        a[t] = 0.0;
        b[t] = 0.0;
    }
}

// fourier.in = jx_bunch(i) - rho_bunch(i) - rho_ion (src/System_3d.cpp:343-347).
__global__ void k_static_src(double *src, const double *jxb, const double *rhob, const double *rho_ion, size_t n,
                             Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) src[t] = (jxb ? jxb[t] : 0.0) - (rhob ? rhob[t] : 0.0) - rho_ion[t];
}

// rho = rho_bunch(i), jx = jx_bunch(i), susceptibility(i) = 0 (src/System_3d.cpp:381-388).
__global__ void k_init_rho(double *rho, double *jx, double *chi, const double *rhob, const double *jxb, size_t n,
                           Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) {
        rho[t] = rhob ? rhob[t] : 0.0;
        jx[t] = jxb ? jxb[t] : 0.0;
        chi[t] = 0.0;
    }
}

// normalize_potential's border mean (src/System_3d.cpp:30-40), summed in the
// reference's order by one thread.
__global__ void k_border_mean(const double *psi, int n1, int n2, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    if (threadIdx.x != 0) return;
    double b = 0.0;
    for (int i = 0; i < n1; ++i) b += psi[(size_t)n2 * i];
    for (int j = 1; j < n2; ++j) b += psi[j];
    b /= (n1 + n2 - 1);
    ctl->border[0] = b;
}

// psi -= border; increase_minimum (src/System_3d.cpp:41-46, 830-841).
__global__ void k_sub_clamp(double *psi, size_t n, double thr, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) {
        double v = psi[t] - ctl->border[0];
        if (v < thr) v = thr;
        psi[t] = v;
    }
}

// d/dxi of the transverse currents (src/System_3d.cpp:489-495).
__global__ void k_djdxi(double *djy, double *djz, const double *jy, const double *jyn, const double *jz,
                        const double *jzn, double dx, size_t n, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t < n) {
        djy[t] = (jy[t] - jyn[t]) / dx;
        djz[t] = (jz[t] - jzn[t]) / dx;
    }
}

// Source of the B_y Helmholtz solve (src/System_3d.cpp:530-536).
__global__ void k_src_by(double *src, const double *by, const double *jx, const double *djz, int n1, int n2,
                         double D, double dz, Ctl *ctl) {
    if (ctl_skip_b(ctl)) return;
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (size_t)n1 * n2) return;
    const int j = (int)(t / n2), k = (int)(t % n2);
    const size_t r = (size_t)n2 * j;
    const double jn = k < n2 - 1 ? jx[r + k + 1] : jx[r];
    src[t] = -D * by[t] - (jn - jx[t]) / dz + djz[t];
}

// Source of the B_z Helmholtz solve (src/System_3d.cpp:555-564).  The last
// row takes bz(i, 0, k), as the reference writes it at :563.
__global__ void k_src_bz(double *src, const double *bz, const double *jx, const double *djy, int n1, int n2,
                         double D, double dy, Ctl *ctl) {
    if (ctl_skip_b(ctl)) return;
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (size_t)n1 * n2) return;
    const int j = (int)(t / n2), k = (int)(t % n2);
    if (j < n1 - 1) {
        src[t] = -D * bz[t] + (jx[t + n2] - jx[t]) / dy - djy[t];
    } else {
        src[t] = -D * bz[k] + (jx[k] - jx[t]) / dy - djy[t];
    }
}

// output_step's fields (src/System_3d.cpp:633-661).
__global__ void k_output(double *ex, double *ey, double *ez, const double *psi, const double *psi_prev,
                         const double *by, const double *bz, int n1, int n2, double dx, double dy, double dz,
                         int with_ex, Ctl *ctl, int final_step) {
    if (!final_step) {
        SKIP_OR_LATCH(ctl);
    } else if (ctl->stop_acc) {
        by = nullptr;  // after a stop, by(0)/bz(0) are still the pass's zeros
        bz = nullptr;
    }
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (size_t)n1 * n2) return;
    const int j = (int)(t / n2), k = (int)(t % n2);
    if (with_ex) ex[t] = (psi[t] - psi_prev[t]) / dx;
    const size_t tj = j < n1 - 1 ? t + n2 : (size_t)k;
    const size_t tk = k < n2 - 1 ? t + 1 : (size_t)n2 * j;
    ey[t] = -(psi[tj] - psi[t]) / dy + (bz ? bz[t] : 0.0);
    ez[t] = -(psi[tk] - psi[t]) / dz - (by ? by[t] : 0.0);
}

// Gaussian <a^2> plane of slice i on the nodes (src/System_3d.cpp:805-816).
__global__ void k_laser_slice(double *asq, LaserGauss L, int i, double dx, double dy, double dz, int n1, int n2) {
    size_t t = (size_t)blockIdx.x * kT + threadIdx.x;
    if (t >= (size_t)n1 * n2) return;
    const int j = (int)(t / n2), k = (int)(t % n2);
    const double x = i * dx, y = j * dy, z = k * dz;
    asq[t] = L.amp * exp(-(x - L.x0) * (x - L.x0) / L.wx / L.wx - (y - L.y0) * (y - L.y0) / L.wy / L.wy -
                         (z - L.z0) * (z - L.z0) / L.wz / L.wz);
}

// --------------------------------------------------------- particle kernels
struct G2 {
    int n1, n2;
    double d1, d2, ly, lz, dx;
};

// deposit(y,z,value,array2d&) with origin (o1,o2), src/array_utils.cpp:85-118.
__device__ __forceinline__ void dep_o(double y, double z, double v, double *a, const G2 &g, double o1, double o2) {
    int j1, j2, k1, k2;
    double yf, zf;
    cic_floor(y, g.d1, o1, g.n1, j1, j2, yf);
    cic_floor(z, g.d2, o2, g.n2, k1, k2, zf);
    scatter4_global(a, g.n2, j1, j2, k1, k2, yf, zf, v);
}

// rho_ion = S(y,z)(-n) (src/System_3d.cpp:318-323).
__global__ void k_dep_ion(const double *y, const double *z, const double *n, double *rho_ion, size_t np, G2 g) {
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p < np) dep_o(y[p], z[p], -n[p], rho_ion, g, 0.0, 0.0);
}

// psi source deposit with the truncating raw-pointer CIC (src/array_utils.cpp:120-142).
__global__ void k_dep_raw(const double *y, const double *z, const double *n, double *src, size_t np, G2 g,
                          Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    int j1, j2, k1, k2;
    double yf, zf;
    cic_trunc(y[p], g.d1, g.n1, j1, j2, yf);
    cic_trunc(z[p], g.d2, g.n2, k1, k2, zf);
    scatter4_global(src, g.n2, j1, j2, k1, k2, yf, zf, -n[p]);
}

// gamma, px (S4: src/System_3d.cpp:368-378; P6 with the mean momentum: :498-508).
__global__ void k_gamma_px(PartFP64 P, const double *asq, const double *psi, size_t np, G2 g, int mean_p,
                           Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    const double y = P.y[p], z = P.z[p];
    const double a = gather_o(y, z, asq, g.n1, g.n2, g.d1, g.d2, 0.0, 0.0);
    const double ps = gather_o(y, z, psi, g.n1, g.n2, g.d1, g.d2, 0.0, 0.0);
    const double py = P.py[p], pz = P.pz[p];
    if (!mean_p) {
        P.gamma[p] = 0.5 * (1 + py * py + pz * pz + a + (1 + ps) * (1 + ps)) / (1 + ps);
        P.px[p] = 0.5 * (1 + py * py + pz * pz + a - (1 + ps) * (1 + ps)) / (1 + ps);
    } else {
        const double pyn = P.pyn[p], pzn = P.pzn[p];
        const double p2 = 0.25 * (py + pyn) * (py + pyn) + 0.25 * (pz + pzn) * (pz + pzn);
        P.gamma[p] = 0.5 * (1 + p2 + a + (1 + ps) * (1 + ps)) / (1 + ps);
        P.px[p] = 0.5 * (1 + p2 + a - (1 + ps) * (1 + ps)) / (1 + ps);
    }
}

// jx, rho, susceptibility deposits (src/System_3d.cpp:390-397, 520-527).
__global__ void k_dep_rho_jx_chi(PartFP64 P, double *jx, double *rho, double *chi, size_t np, G2 g, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    const double n = P.n[p], gm = P.gamma[p];
    const double vx = P.px[p] / gm;
    int j1, j2, k1, k2;
    double yf, zf;
    cic_floor(P.y[p], g.d1, 0.0, g.n1, j1, j2, yf);
    cic_floor(P.z[p], g.d2, 0.0, g.n2, k1, k2, zf);
    scatter4_global(jx, g.n2, j1, j2, k1, k2, yf, zf, n * vx / (1 - vx));
    scatter4_global(rho, g.n2, j1, j2, k1, k2, yf, zf, n / (1 - vx));
    scatter4_global(chi, g.n2, j1, j2, k1, k2, yf, zf, -n / (1 - vx) / gm);
}

// Momentum predictor (src/System_3d.cpp:416-433).
__global__ void k_predictor(PartFP64 P, const double *asq, const double *psi, const double *by, const double *bz,
                            size_t np, G2 g, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    const double y = P.y[p], z = P.z[p];
    const double ps = gather_o(y, z, psi, g.n1, g.n2, g.d1, g.d2, 0.0, 0.0);
    const double day = gather_yder(y, z, asq, g.n1, g.n2, g.d1, g.d2);
    const double daz = gather_zder(y, z, asq, g.n1, g.n2, g.d1, g.d2);
    const double dpy = gather_yder(y, z, psi, g.n1, g.n2, g.d1, g.d2);
    const double dpz = gather_zder(y, z, psi, g.n1, g.n2, g.d1, g.d2);
    const double byp = gather_o(y, z, by, g.n1, g.n2, g.d1, g.d2, 0.0, 0.5 * g.d2);
    const double bzp = gather_o(y, z, bz, g.n1, g.n2, g.d1, g.d2, 0.5 * g.d1, 0.0);
    const double dx = g.dx, gm = P.gamma[p];
    double pyn = P.py[p] - dx * 0.5 * day / (1 + ps);
    pyn += dx * gm * dpy / (1 + ps);
    pyn -= dx * bzp;
    double pzn = P.pz[p] - dx * 0.5 * daz / (1 + ps);
    pzn += dx * gm * dpz / (1 + ps);
    pzn += dx * byp;
    P.pyn[p] = pyn;
    P.pzn[p] = pzn;
}

// Half-step coordinates (src/System_3d.cpp:436-443).
__global__ void k_half(PartFP64 P, const double *psi, size_t np, G2 g, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    const double y = P.y[p], z = P.z[p];
    const double ps = gather_o(y, z, psi, g.n1, g.n2, g.d1, g.d2, 0.0, 0.0);
    double ym = y + 0.5 * g.dx * P.pyn[p] / (1 + ps);
    double zm = z + 0.5 * g.dx * P.pzn[p] / (1 + ps);
    wrap_coords(ym, zm, g.ly, g.lz);
    P.ym[p] = ym;
    P.zm[p] = zm;
}

// Transverse current at the half-step positions (src/System_3d.cpp:480-486).
__global__ void k_dep_jperp(PartFP64 P, const double *psim, double *jyn, double *jzn, size_t np, G2 g, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    const double ym = P.ym[p], zm = P.zm[p], n = P.n[p];
    const double ps = gather_o(ym, zm, psim, g.n1, g.n2, g.d1, g.d2, 0.0, 0.0);
    dep_o(ym, zm, n * P.pyn[p] / (1 + ps), jyn, g, 0.5 * g.d1, 0.0);
    dep_o(ym, zm, n * P.pzn[p] / (1 + ps), jzn, g, 0.0, 0.5 * g.d2);
}

// p = p* (src/System_3d.cpp:593-598).
__global__ void k_commit(PartFP64 P, size_t np, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    P.py[p] = P.pyn[p];
    P.pz[p] = P.pzn[p];
}

// Coordinate advance with psi_middle at the old position (src/System_3d.cpp:601-608).
__global__ void k_advance(PartFP64 P, const double *psim, size_t np, G2 g, Ctl *ctl) {
    SKIP_OR_LATCH(ctl);
    size_t p = (size_t)blockIdx.x * kT + threadIdx.x;
    if (p >= np) return;
    double y = P.y[p], z = P.z[p];
    const double ps = gather_o(y, z, psim, g.n1, g.n2, g.d1, g.d2, 0.0, 0.0);
    y += g.dx * P.py[p] / (1 + ps);
    z += g.dx * P.pz[p] / (1 + ps);
    wrap_coords(y, z, g.ly, g.lz);
    P.y[p] = y;
    P.z[p] = z;
}

// ------------------------------------------------------------------ driver
#define LAUNCH(kern, blocks, ...)                                              \
    do {                                                                       \
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kT), 0, s, __VA_ARGS__);   \
        QP_HIP_CHECK(hipGetLastError());                                       \
    } while (0)

int wake_upload_fp64(Wake *w) {
    const size_t np = w->np();
    hipStream_t s = w->stream();
    PartFP64 &P = w->pf;
    if (w->cap < np || !P.y) {
        double **arrs[] = {&P.y, &P.z, &P.py, &P.pz, &P.n, &P.gamma, &P.px, &P.ym, &P.zm, &P.pyn, &P.pzn,
                           &w->init64[0], &w->init64[1], &w->init64[2], &w->init64[3], &w->init64[4]};
        for (double **a : arrs) {
            if (*a) (void)hipFree(*a);
            QP_HIP_CHECK(hipMalloc(a, sizeof(double) * (np ? np : 1)));
        }
        w->cap = np;
        w->particles_dirty = true;
    }
    const size_t b = sizeof(double) * np;
    if (w->particles_dirty) {
        const WakeMember &m0 = w->mem[0];
        const std::vector<double> *src[5] = {&m0.h_y, &m0.h_z, &m0.h_py, &m0.h_pz, &m0.h_w};
        for (int i = 0; i < 5; ++i)
            QP_HIP_CHECK(hipMemcpy(w->init64[i], src[i]->data(), b, hipMemcpyHostToDevice));
        w->particles_dirty = false;
    }
    double *dst[5] = {P.y, P.z, P.py, P.pz, P.n};
    for (int i = 0; i < 5; ++i) QP_HIP_CHECK(hipMemcpyAsync(dst[i], w->init64[i], b, hipMemcpyDeviceToDevice, s));
    // init_wake_particles: py_next = pz_next = 0, gamma = 0 (src/System_3d.cpp:740-742)
    QP_HIP_CHECK(hipMemsetAsync(P.pyn, 0, b, s));
    QP_HIP_CHECK(hipMemsetAsync(P.pzn, 0, b, s));
    QP_HIP_CHECK(hipMemsetAsync(P.gamma, 0, b, s));
    QP_HIP_CHECK(hipMemsetAsync(P.px, 0, b, s));
    QP_HIP_CHECK(hipMemsetAsync(P.ym, 0, b, s));
    QP_HIP_CHECK(hipMemsetAsync(P.zm, 0, b, s));
    return QP_OK;
}

int wake_download_fp64(Wake *w, double *y, double *z, double *py, double *pz) {
    const size_t b = sizeof(double) * w->np();
    QP_HIP_CHECK(hipStreamSynchronize(w->stream()));
    if (y) QP_HIP_CHECK(hipMemcpy(y, w->pf.y, b, hipMemcpyDeviceToHost));
    if (z) QP_HIP_CHECK(hipMemcpy(z, w->pf.z, b, hipMemcpyDeviceToHost));
    if (py) QP_HIP_CHECK(hipMemcpy(py, w->pf.py, b, hipMemcpyDeviceToHost));
    if (pz) QP_HIP_CHECK(hipMemcpy(pz, w->pf.pz, b, hipMemcpyDeviceToHost));
    return QP_OK;
}

static int solve_plain(Wake *w, const double *src, double *dst, double D, bool bsolve, int slice) {
    hipStream_t s = w->stream();
    const Solver2d &S = w->S;
    prof_begin(w, "solve");
    if (!bsolve) {
        QP_TRY(solve_rows_fwd_plain_ctl(S, src, S.spec[0], s, w->ctl, 1));
        QP_TRY(solve_cols(S, D, 1, s, w->ctl, 1));
        QP_TRY(solve_rows_inv_plain_ctl(S, S.spec[0], dst, s, w->ctl));
    } else {
        QP_TRY(solve_rows_fwd_plain_ctl(S, src, S.spec[0], s, w->ctl, 2));
        QP_TRY(solve_cols(S, D, 1, s, w->ctl, 2));
        QP_TRY(solve_rows_inv_bcheck(S, S.spec[0], dst, w->p.b_threshold, w->ctl, slice, s));
    }
    prof_end(w, "solve");
    return QP_OK;
}

int wake_run_fp64(Wake *w, qp_slice_cb cb, void *user, int *stop_slice) {
    const qp_wake_params &p = w->p;
    hipStream_t s = w->stream();
    const size_t nn = w->nn(), np = w->np();
    const int n1 = p.ny, n2 = p.nz;
    const G2 g{n1, n2, p.dy, p.dz, p.ly, p.lz, p.dx};
    const double thr = p.psi_threshold - 1;
    const int gb = nblk(nn), pb = nblk(np);
    Ctl *ctl = w->ctl;

    // Upload first: it allocates the particle arrays on the first pass.
    QP_TRY(wake_upload_fp64(w));
    const PartFP64 P = w->pf;
    if (np && !(P.y && P.z && P.py && P.pz && P.n && P.gamma && P.px && P.ym && P.zm && P.pyn && P.pzn)) {
        set_error("wake_run_fp64: particle arrays not allocated");
        return QP_ERR_STATE;
    }
    QP_HIP_CHECK(hipMemsetAsync(ctl, 0, sizeof(Ctl), s));
    // zero psi, by, bz, rho_ion, jy, jz (src/System_3d.cpp:265-316)
    double *zero_list[] = {w->psi, w->psi_prev, w->by, w->bz, w->rho_ion, w->jy, w->jz, w->jyn, w->jzn,
                           w->psim, w->rho, w->jx, w->chi, w->ex, w->ey, w->ez, w->djy, w->djz};
    for (double *a : zero_list) QP_HIP_CHECK(hipMemsetAsync(a, 0, sizeof(double) * nn, s));
    QP_HIP_CHECK(hipEventRecord(w->ev0, s));
    if (np) LAUNCH(k_dep_ion, pb, P.y, P.z, P.n, w->rho_ion, np, g);
    if (w->bunch) QP_TRY(bunch_clear_fields(w->bunch, s));

    int delivered_stop = 0;
    const int i_end = (p.slice_limit > 0 && p.slice_limit < p.nx) ? p.nx - p.slice_limit : 0;
    for (int i = p.nx - 1; i >= i_end; --i) {
        const double *asq;
        QP_TRY(wake_asq_slice(w, i, &asq));
        if (w->bunch) QP_TRY(bunch_slice_sources(w->bunch, ctl, -1, i, s));
        const double *rb = wake_rho_b(w, i), *jb = wake_jx_b(w, i);
        // S1-S3
        LAUNCH(k_copy, gb, w->psi_prev, w->psi, nn, ctl);
        LAUNCH(k_static_src, gb, w->src, jb, rb, w->rho_ion, nn, ctl);
        prof_begin(w, "push");
        if (np) LAUNCH(k_dep_raw, pb, P.y, P.z, P.n, w->src, np, g, ctl);
        prof_end(w, "push");
        QP_TRY(solve_plain(w, w->src, w->psi, 0.0, false, i));
        LAUNCH(k_border_mean, 1, w->psi, n1, n2, ctl);
        LAUNCH(k_sub_clamp, gb, w->psi, nn, thr, ctl);
        // S4-S5
        prof_begin(w, "push");
        if (np) LAUNCH(k_gamma_px, pb, P, asq, w->psi, np, g, 0, ctl);
        prof_end(w, "push");
        LAUNCH(k_init_rho, gb, w->rho, w->jx, w->chi, rb, jb, nn, ctl);
        prof_begin(w, "push");
        if (np) LAUNCH(k_dep_rho_jx_chi, pb, P, w->jx, w->rho, w->chi, np, g, ctl);
        prof_end(w, "push");
        // S6: by(i) = by(i+1) is the carried buffer
        if (i == 0) {
            if (w->env) LAUNCH(k_copy, gb, w->chi_store(i), (const double *)w->chi, nn, ctl);
            break;
        }
        for (int it = 0; it < p.K; ++it) {
            prof_begin(w, "push");
            if (np) LAUNCH(k_predictor, pb, P, asq, w->psi, w->by, w->bz, np, g, ctl);
            if (np) LAUNCH(k_half, pb, P, w->psi, np, g, ctl);
            prof_end(w, "push");
            LAUNCH(k_static_src, gb, w->src, jb, rb, w->rho_ion, nn, ctl);
            prof_begin(w, "push");
            if (np) LAUNCH(k_dep_raw, pb, P.ym, P.zm, P.n, w->src, np, g, ctl);
            prof_end(w, "push");
            QP_TRY(solve_plain(w, w->src, w->psim, 0.0, false, i));
            LAUNCH(k_border_mean, 1, w->psim, n1, n2, ctl);
            LAUNCH(k_sub_clamp, gb, w->psim, nn, thr, ctl);
            LAUNCH(k_zero2, gb, w->jyn, w->jzn, nn, ctl);
            prof_begin(w, "push");
            if (np) LAUNCH(k_dep_jperp, pb, P, w->psim, w->jyn, w->jzn, np, g, ctl);
            prof_end(w, "push");
            LAUNCH(k_djdxi, gb, w->djy, w->djz, w->jy, w->jyn, w->jz, w->jzn, p.dx, nn, ctl);
            prof_begin(w, "push");
            if (np) LAUNCH(k_gamma_px, pb, P, asq, w->psi, np, g, 1, ctl);
            prof_end(w, "push");
            LAUNCH(k_init_rho, gb, w->rho, w->jx, w->chi, rb, jb, nn, ctl);
            prof_begin(w, "push");
            if (np) LAUNCH(k_dep_rho_jx_chi, pb, P, w->jx, w->rho, w->chi, np, g, ctl);
            prof_end(w, "push");
            LAUNCH(k_src_by, gb, w->src, w->by, w->jx, w->djz, n1, n2, p.D, p.dz, ctl);
            QP_TRY(solve_plain(w, w->src, w->by, p.D, true, i));
            LAUNCH(k_src_bz, gb, w->src, w->bz, w->jx, w->djy, n1, n2, p.D, p.dy, ctl);
            QP_TRY(solve_plain(w, w->src, w->bz, p.D, true, i));
        }
        if (w->env) LAUNCH(k_copy, gb, w->chi_store(i), (const double *)w->chi, nn, ctl);
        prof_begin(w, "push");
        if (np) LAUNCH(k_commit, pb, P, np, ctl);
        if (np) LAUNCH(k_advance, pb, P, w->psim, np, g, ctl);
        prof_end(w, "push");
        LAUNCH(k_output, gb, w->ex, w->ey, w->ez, w->psi, w->psi_prev, w->by, w->bz, n1, n2, p.dx, p.dy, p.dz,
               i < p.nx - 1 ? 1 : 0, ctl, 0);
        if (w->bunch) QP_TRY(bunch_slice_gather(w->bunch, w->ex, w->ey, w->ez, w->by, w->bz, ctl, -1, i, s));
        if (cb) {
            QP_HIP_CHECK(hipStreamSynchronize(s));
            Ctl hc;
            QP_HIP_CHECK(hipMemcpy(&hc, ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
            if (hc.stop_acc) {
                QP_TRY(deliver_slice(w, cb, user, hc.stop_slice, QP_SLICE_STOPPED, w->psi, w->rho, w->jx, w->jy,
                                     w->jz, w->ex, w->ey, w->ez, w->by, w->bz, w->chi));
                delivered_stop = 1;
                break;
            }
            QP_TRY(deliver_slice(w, cb, user, i, QP_SLICE_NORMAL, w->psi, w->rho, w->jx, w->jy, w->jz, w->ex,
                                 w->ey, w->ez, w->by, w->bz, w->chi));
        }
        LAUNCH(k_copy, gb, w->jy, w->jyn, nn, ctl);
        LAUNCH(k_copy, gb, w->jz, w->jzn, nn, ctl);
    }
    if (i_end == 0) {
        // output_step(0) after the loop (src/System_3d.cpp:622)
        LAUNCH(k_output, gb, w->ex, w->ey, w->ez, w->psi, w->psi_prev, w->by, w->bz, n1, n2, p.dx, p.dy, p.dz,
               p.nx > 1 ? 1 : 0, ctl, 1);
        // bunch bucket 0: slice 0's fields are final only after output_step(0)
        if (w->bunch) QP_TRY(bunch_slice_gather(w->bunch, w->ex, w->ey, w->ez, w->by, w->bz, ctl, -1, 0, s));
        // the envelope step after the pass (the plain path keeps the launches in order)
        if (w->env_step()) QP_TRY(envelope_step(w->env, w->env_dtau, w->env_chi_it, w->p.use_graph, s));
    }
    QP_HIP_CHECK(hipEventRecord(w->ev1, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    QP_HIP_CHECK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    w->last_ms = ms;
    Ctl hc;
    QP_HIP_CHECK(hipMemcpy(&hc, ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (hc.stop_acc && cb && !delivered_stop) {
        QP_TRY(deliver_slice(w, cb, user, hc.stop_slice, QP_SLICE_STOPPED, w->psi, w->rho, w->jx, w->jy, w->jz,
                             w->ex, w->ey, w->ez, w->by, w->bz, w->chi));
    }
    if (cb && i_end == 0) {
        const double *bz0 = hc.stop_acc ? w->zero2 : w->bz;
        const double *by0 = hc.stop_acc ? w->zero2 : w->by;
        const double *chi0 = hc.stop_acc ? w->zero2 : w->chi;
        QP_TRY(deliver_slice(w, cb, user, 0, QP_SLICE_FINAL, w->psi, w->rho, w->jx, w->jy, w->jz, w->ex, w->ey,
                             w->ez, by0, bz0, chi0));
    }
    if (stop_slice) *stop_slice = hc.stop_acc ? hc.stop_slice : -1;
    return hc.stop_acc ? QP_STOPPED : QP_OK;
}

}  // namespace qp

namespace qp {

int wake_asq_slice(Wake *w, int i, const double **asq) {
    const qp_wake_params &p = w->p;
    if (w->asq3()) {
        *asq = w->asq3() + (size_t)i * w->nn();
        return QP_OK;
    }
    if (w->mem[0].laser_set) {
        hipStream_t s = w->stream();
        LAUNCH(k_laser_slice, nblk(w->nn()), w->asq2, w->mem[0].laser, i, p.dx, p.dy, p.dz, p.ny, p.nz);
        *asq = w->asq2;
        return QP_OK;
    }
    *asq = w->zero2;  // no [laser] block: a_sqr stays zero (src/System_parameters.h:22)
    return QP_OK;
}

}  // namespace qp
