// qp_wake_mixed.hip -- the fused mixed-precision slice loop.
//
// Same algorithm as System_3d::solve_wakefield (src/System_3d.cpp:238-627),
// reorganised around its dependency structure:
//   slice i:  A   commit p = p*, advance with psi_middle of slice i+1 (:593-608),
//                 deposit the psi source at the new positions (:342-353)
//             psi solve with the source assembled in the row pass and
//                 normalize_potential + clamp applied in the last pass (:355-365)
//             per iteration k = 1..K:
//               B  gamma (S4 for k = 1, else P6 recomputed from p, p*_{k-1}),
//                  predictor P1, half step P2, psi_middle source deposit P3
//               psi_middle solve (:459-469)
//               C  half-step position recomputed from the same inputs, j_perp*
//                  deposit P4, gamma/px P6, rho/jx/chi deposits P7
//               B_y and B_z Helmholtz solves in one set of launches, sources
//                  (P5, P8, P9) assembled in the row pass, stop rule in the
//                  last pass (:529-580)
//             output fields e_x, e_y, e_z (:629-661)
// The S5 deposits of a slice with K >= 1 are overwritten by P7 before anyone
// reads them (:380-397 vs :510-527) and are not recomputed; slice 0 (and
// K = 0) run S4 + S5 instead of the iterations (:409-411).
#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>
#include <vector>

#include "qp_mixed_kernels.h"
#include "qp_wave_solver.h"

namespace qp {

static inline int pblocks(int np) { return np > 0 ? (np + kPT - 1) / kPT : 1; }

#define MX_SKIP(ctl)              \
    if (ctl_skip(ctl)) {          \
        ctl_latch(ctl);           \
        return;                   \
    }

// ------------------------------------------------------------------ A phase
// Commit and advance (src/System_3d.cpp:593-608), then deposit -n at the new
// position into the psi source accumulator (:349-353).  The grid part
// evaluates the laser plane of slice i and, for S-phase slices, initialises
// rho/jx/chi (:381-388).
__global__ __launch_bounds__(kPT) void k_mx_A(MxPart P, MxGeom g, MxSliceIn si, const double *psim, double *src_acc,
                                              double *asq2, LaserGauss L, int laser, double *rho, double *jx,
                                              double *chi, int init_rho, int advance, Ctl *ctl) {
    MX_SKIP(ctl);
    __shared__ double lds[kWinCap];
    __shared__ int sh[4 * (kPT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2;
    const int slice = mx_slice(ctl, si);
    if (laser || init_rho) {
        const double *rb = si.rho_b3 ? si.rho_b3 + (size_t)slice * nn : nullptr;
        const double *jb = si.jx_b3 ? si.jx_b3 + (size_t)slice * nn : nullptr;
        for (size_t t = (size_t)blockIdx.x * kPT + threadIdx.x; t < nn; t += (size_t)gridDim.x * kPT) {
            if (laser) {
                const int j = (int)(t / n2), k = (int)(t % n2);
                const double x = slice * g.dx, y = j * g.d1, z = k * g.d2;
                asq2[t] = L.amp * exp(-(x - L.x0) * (x - L.x0) / L.wx / L.wx - (y - L.y0) * (y - L.y0) / L.wy / L.wy -
                                      (z - L.z0) * (z - L.z0) / L.wz / L.wz);
            }
            if (init_rho) {
                rho[t] = rb ? rb[t] : 0.0;
                jx[t] = jb ? jb[t] : 0.0;
                chi[t] = 0.0;
            }
        }
    }
    const int p = blockIdx.x * kPT + threadIdx.x;
    const bool valid = p < P.np;
    Pos q{0, 0, 0.0, 0.0};
    double w = 0.0;
    if (valid) {
        q = load_pos(P, p);
        w = P.w[p];
        if (advance) {
            // commit p <- p* (advance 2: K = 0, where p* keeps its initial value 0)
            const double pyv = advance == 2 ? 0.0 : P.py[p] + (double)P.dpy[p];
            const double pzv = advance == 2 ? 0.0 : P.pz[p] + (double)P.dpz[p];
            P.py[p] = pyv;
            P.pz[p] = pzv;
            const double ps = mx_gather(psim, q, n1, n2);
            const double s = g.dx / (1 + ps);
            double Y = (q.j + q.fy) + (s * pyv) * g.inv_d1;
            double Z = (q.k + q.fz) + (s * pzv) * g.inv_d2;
            Y = wrap_cells(Y, n1);
            Z = wrap_cells(Z, n2);
            int j, k;
            uint32_t oy, oz;
            split_coord(Y, n1, j, oy);
            split_coord(Z, n2, k, oz);
            P.cell[p] = ((uint32_t)j << 16) | (uint32_t)k;
            P.oy[p] = oy;
            P.oz[p] = oz;
            q.j = j;
            q.k = k;
            q.fy = (double)oy * kInvTwo32;
            q.fz = (double)oz * kInvTwo32;
        }
    }
    if (threadIdx.x == 0) {
        sh[4 * (kPT / 64)] = q.j;
        sh[4 * (kPT / 64) + 1] = q.k;
    }
    __syncthreads();
    const int jref = sh[4 * (kPT / 64)], kref = sh[4 * (kPT / 64) + 1];
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) range_add(q.j, q.k, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    const Win win = block_window(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_zero(lds, 1, win);
    __syncthreads();
    {
        int j1, j2, k1, k2;
        double yf, zf;
        ax_node(q.j, q.fy, n1, j1, j2, yf);
        ax_node(q.k, q.fz, n2, k1, k2, zf);
        double a[4];
        cic_w(-w, yf, zf, a);
        bool act = valid;
        quad_fold<4>(j1 * n2 + k1, act, a);
        if (act) dep4w(lds, win, src_acc, n1, n2, j1, j2, k1, k2, a[0], a[1], a[2], a[3]);
    }
    __syncthreads();
    double *gl[1] = {src_acc};
    win_flush(lds, gl, 1, win, n1, n2);
}

// ------------------------------------------------------------------ B phase
// gamma = (1 + p^2 + a + (1+psi)^2) / (2 (1+psi)) (src/System_3d.cpp:373-374), with 1/(1+psi) given.
__device__ __forceinline__ double gamma_of(double p2, double a, double ps, double inv1ps) {
    return 0.5 * (1 + p2 + a + (1 + ps) * (1 + ps)) * inv1ps;
}

__global__ __launch_bounds__(kPT) void k_mx_B(MxPart P, MxGeom g, MxSliceIn si, const double *psi, const double *by,
                                              const double *bz, double *src_acc, double *rho, double *jx,
                                              double *chi, double *jyn, double *jzn, int first_iter, Ctl *ctl) {
    MX_SKIP(ctl);
    __shared__ double lds[kWinCap];
    __shared__ int sh[4 * (kPT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2;
    const int slice = mx_slice(ctl, si);
    {  // rho = rho_bunch(i), jx = jx_bunch(i), chi = 0, jy* = jz* = 0 before the C phase (:473-478, :511-518)
        const double *rb = si.rho_b3 ? si.rho_b3 + (size_t)slice * nn : nullptr;
        const double *jb = si.jx_b3 ? si.jx_b3 + (size_t)slice * nn : nullptr;
        for (size_t t = (size_t)blockIdx.x * kPT + threadIdx.x; t < nn; t += (size_t)gridDim.x * kPT) {
            rho[t] = rb ? rb[t] : 0.0;
            jx[t] = jb ? jb[t] : 0.0;
            chi[t] = 0.0;
            jyn[t] = 0.0;
            jzn[t] = 0.0;
        }
    }
    const double *asq = mx_asq(ctl, si, nn);
    const int p = blockIdx.x * kPT + threadIdx.x;
    const bool valid = p < P.np;
    Pos qm{0, 0, 0.0, 0.0};
    double w = 0.0;
    if (valid) {
        const Pos q = load_pos(P, p);
        w = P.w[p];
        const double py = P.py[p], pz = P.pz[p];
        const double a = mx_gather(asq, q, n1, n2);
        const double ps = mx_gather(psi, q, n1, n2);
        const double inv1ps = 1.0 / (1 + ps);  // the FP64 1/(1+psi) factor
        double gm;
        if (first_iter) {
            gm = gamma_of(py * py + pz * pz, a, ps, inv1ps);  // S4 (:373-374)
        } else {
            const double pys = py + (double)P.dpy[p], pzs = pz + (double)P.dpz[p];  // P6 of the previous iteration (:503-505)
            gm = gamma_of(0.25 * (py + pys) * (py + pys) + 0.25 * (pz + pzs) * (pz + pzs), a, ps, inv1ps);
        }
        const double day = mx_yder(asq, q, n1, n2, g.inv_d1);
        const double daz = mx_zder(asq, q, n1, n2, g.inv_d2);
        const double dpy = mx_yder(psi, q, n1, n2, g.inv_d1);
        const double dpz = mx_zder(psi, q, n1, n2, g.inv_d2);
        const double byp = mx_gather_hz(by, q, n1, n2);
        const double bzp = mx_gather_hy(bz, q, n1, n2);
        const double dx = g.dx;
        const double s = dx * inv1ps;
        double pyn = py - s * 0.5 * day;  // P1 (:427-432)
        pyn += s * gm * dpy;
        pyn -= dx * bzp;
        double pzn = pz - s * 0.5 * daz;
        pzn += s * gm * dpz;
        pzn += dx * byp;
        const float dyf = (float)(pyn - py), dzf = (float)(pzn - pz);
        P.dpy[p] = dyf;
        P.dpz[p] = dzf;
        // P2 (:439-442) with the stored p* = p + increment
        const double pyst = py + (double)dyf, pzst = pz + (double)dzf;
        double Ym = (q.j + q.fy) + (0.5 * s * pyst) * g.inv_d1;
        double Zm = (q.k + q.fz) + (0.5 * s * pzst) * g.inv_d2;
        qm = pos_from_cells(wrap_cells(Ym, n1), wrap_cells(Zm, n2), n1, n2);
    }
    if (threadIdx.x == 0) {
        sh[4 * (kPT / 64)] = qm.j;
        sh[4 * (kPT / 64) + 1] = qm.k;
    }
    __syncthreads();
    const int jref = sh[4 * (kPT / 64)], kref = sh[4 * (kPT / 64) + 1];
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) range_add(qm.j, qm.k, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    const Win win = block_window(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_zero(lds, 1, win);
    __syncthreads();
    {  // P3 (:453-457)
        int j1, j2, k1, k2;
        double yf, zf;
        ax_node(qm.j, qm.fy, n1, j1, j2, yf);
        ax_node(qm.k, qm.fz, n2, k1, k2, zf);
        double a[4];
        cic_w(-w, yf, zf, a);
        bool act = valid;
        quad_fold<4>(j1 * n2 + k1, act, a);
        if (act) dep4w(lds, win, src_acc, n1, n2, j1, j2, k1, k2, a[0], a[1], a[2], a[3]);
    }
    __syncthreads();
    double *gl[1] = {src_acc};
    win_flush(lds, gl, 1, win, n1, n2);
}

// ------------------------------------------------------------------ C phase
// ngr = 5 grids: jy*, jz*, rho, jx, chi.
__global__ __launch_bounds__(kPT) void k_mx_C(MxPart P, MxGeom g, MxSliceIn si, const double *psi,
                                              const double *psim, double *rho, double *jx, double *chi,
                                              double *jyn, double *jzn, Ctl *ctl) {
    MX_SKIP(ctl);
    extern __shared__ double ldsd[];
    double *lds = ldsd;
    __shared__ int sh[4 * (kPT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2;
    const double *asq = mx_asq(ctl, si, nn);
    const int p = blockIdx.x * kPT + threadIdx.x;
    const bool valid = p < P.np;
    Pos q{0, 0, 0.0, 0.0}, qm{0, 0, 0.0, 0.0};
    double vjy = 0, vjz = 0, vjx = 0, vrho = 0, vchi = 0;
    if (valid) {
        q = load_pos(P, p);
        const double w = P.w[p];
        const double py = P.py[p], pz = P.pz[p];
        const double pys = py + (double)P.dpy[p], pzs = pz + (double)P.dpz[p];
        const double ps = mx_gather(psi, q, n1, n2);
        const double inv1ps = 1.0 / (1 + ps);
        const double s = g.dx * inv1ps;
        double Ym = (q.j + q.fy) + (0.5 * s * pys) * g.inv_d1;
        double Zm = (q.k + q.fz) + (0.5 * s * pzs) * g.inv_d2;
        qm = pos_from_cells(wrap_cells(Ym, n1), wrap_cells(Zm, n2), n1, n2);
        const double psm = mx_gather(psim, qm, n1, n2);  // P4 (:483-485)
        const double wm = w / (1 + psm);
        vjy = wm * pys;
        vjz = wm * pzs;
        const double a = mx_gather(asq, q, n1, n2);  // P6 (:501-506)
        const double p2 = 0.25 * (py + pys) * (py + pys) + 0.25 * (pz + pzs) * (pz + pzs);
        const double gm = gamma_of(p2, a, ps, inv1ps);
        const double px = 0.5 * (1 + p2 + a - (1 + ps) * (1 + ps)) * inv1ps;
        // P7 (:523-526): vx = px/gamma, 1/(1-vx) = gamma/(gamma-px) = gamma/(1+psi)
        const double vx = px / gm;
        const double inv1mvx = 1.0 / (1 - vx);
        vjx = w * vx * inv1mvx;
        vrho = w * inv1mvx;
        vchi = -w * inv1mvx / gm;
    }
    if (threadIdx.x == 0) {
        sh[4 * (kPT / 64)] = q.j;
        sh[4 * (kPT / 64) + 1] = q.k;
    }
    __syncthreads();
    const int jref = sh[4 * (kPT / 64)], kref = sh[4 * (kPT / 64) + 1];
    int ya1, ya2, za1, za2, yb1, yb2, zb1, zb2, yc1, yc2, zc1, zc2;
    double yaf, zaf, ybf, zbf, ycf, zcf;
    ax_half(qm.j, qm.fy, n1, ya1, ya2, yaf);  // jy* on the (dy/2, 0) grid
    ax_node(qm.k, qm.fz, n2, za1, za2, zaf);
    ax_node(qm.j, qm.fy, n1, yb1, yb2, ybf);  // jz* on the (0, dz/2) grid
    ax_half(qm.k, qm.fz, n2, zb1, zb2, zbf);
    ax_node(q.j, q.fy, n1, yc1, yc2, ycf);    // rho, jx, chi on the nodes
    ax_node(q.k, q.fz, n2, zc1, zc2, zcf);
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) {
        range_add(ya1, za1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
        range_add(yb1, zb1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
        range_add(yc1, zc1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    }
    const Win win = block_window(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_zero(lds, 5, win);
    __syncthreads();
    {
        double aa[4], ab[4], ac[12];
        cic_w(vjy, yaf, zaf, aa);
        cic_w(vjz, ybf, zbf, ab);
        cic_w(vrho, ycf, zcf, ac);
        cic_w(vjx, ycf, zcf, ac + 4);
        cic_w(vchi, ycf, zcf, ac + 8);
        bool fa = valid, fb = valid, fc = valid;
        quad_fold<4>(ya1 * n2 + za1, fa, aa);
        quad_fold<4>(yb1 * n2 + zb1, fb, ab);
        quad_fold<12>(yc1 * n2 + zc1, fc, ac);
        if (fa) dep4w(lds, win, jyn, n1, n2, ya1, ya2, za1, za2, aa[0], aa[1], aa[2], aa[3]);
        if (fb) dep4w(lds + kWinCap, win, jzn, n1, n2, yb1, yb2, zb1, zb2, ab[0], ab[1], ab[2], ab[3]);
        if (fc) {
            dep4w(lds + 2 * kWinCap, win, rho, n1, n2, yc1, yc2, zc1, zc2, ac[0], ac[1], ac[2], ac[3]);
            dep4w(lds + 3 * kWinCap, win, jx, n1, n2, yc1, yc2, zc1, zc2, ac[4], ac[5], ac[6], ac[7]);
            dep4w(lds + 4 * kWinCap, win, chi, n1, n2, yc1, yc2, zc1, zc2, ac[8], ac[9], ac[10], ac[11]);
        }
    }
    __syncthreads();
    double *gl[5] = {jyn, jzn, rho, jx, chi};
    win_flush(lds, gl, 5, win, n1, n2);
}

// -------------------------------------------------- S phase (slice 0, K = 0)
// gamma/px (S4, :368-378) and the rho/jx/chi deposits (S5, :390-397).
__global__ __launch_bounds__(kPT) void k_mx_S(MxPart P, MxGeom g, MxSliceIn si, const double *psi, double *rho,
                                              double *jx, double *chi, Ctl *ctl) {
    MX_SKIP(ctl);
    extern __shared__ double ldsd[];
    double *lds = ldsd;
    __shared__ int sh[4 * (kPT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2;
    const double *asq = mx_asq(ctl, si, nn);
    const int p = blockIdx.x * kPT + threadIdx.x;
    const bool valid = p < P.np;
    Pos q{0, 0, 0.0, 0.0};
    double vjx = 0, vrho = 0, vchi = 0;
    if (valid) {
        q = load_pos(P, p);
        const double w = P.w[p];
        const double py = P.py[p], pz = P.pz[p];
        const double ps = mx_gather(psi, q, n1, n2);
        const double a = mx_gather(asq, q, n1, n2);
        const double inv1ps = 1.0 / (1 + ps);
        const double gm = gamma_of(py * py + pz * pz, a, ps, inv1ps);
        const double px = 0.5 * (1 + py * py + pz * pz + a - (1 + ps) * (1 + ps)) * inv1ps;
        const double vx = px / gm;
        const double inv1mvx = 1.0 / (1 - vx);
        vjx = w * vx * inv1mvx;
        vrho = w * inv1mvx;
        vchi = -w * inv1mvx / gm;
    }
    if (threadIdx.x == 0) {
        sh[4 * (kPT / 64)] = q.j;
        sh[4 * (kPT / 64) + 1] = q.k;
    }
    __syncthreads();
    const int jref = sh[4 * (kPT / 64)], kref = sh[4 * (kPT / 64) + 1];
    int y1, y2, z1, z2;
    double yf, zf;
    ax_node(q.j, q.fy, n1, y1, y2, yf);
    ax_node(q.k, q.fz, n2, z1, z2, zf);
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) range_add(y1, z1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    const Win win = block_window(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_zero(lds, 3, win);
    __syncthreads();
    {
        double ac[12];
        cic_w(vrho, yf, zf, ac);
        cic_w(vjx, yf, zf, ac + 4);
        cic_w(vchi, yf, zf, ac + 8);
        bool act = valid;
        quad_fold<12>(y1 * n2 + z1, act, ac);
        if (act) {
            dep4w(lds, win, rho, n1, n2, y1, y2, z1, z2, ac[0], ac[1], ac[2], ac[3]);
            dep4w(lds + kWinCap, win, jx, n1, n2, y1, y2, z1, z2, ac[4], ac[5], ac[6], ac[7]);
            dep4w(lds + 2 * kWinCap, win, chi, n1, n2, y1, y2, z1, z2, ac[8], ac[9], ac[10], ac[11]);
        }
    }
    __syncthreads();
    double *gl[3] = {rho, jx, chi};
    win_flush(lds, gl, 3, win, n1, n2);
}

// rho_ion = S(y,z)(-n) (src/System_3d.cpp:318-323), direct atomics (once per pass).
__global__ __launch_bounds__(kPT) void k_mx_ion(MxPart P, MxGeom g, double *rho_ion) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    const Pos q = load_pos(P, p);
    int j1, j2, k1, k2;
    double yf, zf;
    ax_node(q.j, q.fy, g.n1, j1, j2, yf);
    ax_node(q.k, q.fz, g.n2, k1, k2, zf);
    scatter4_global(rho_ion, g.n2, j1, j2, k1, k2, yf, zf, -(double)P.w[p]);
}

// ------------------------------------------------------------- output step
// e_x, e_y, e_z of slice i (src/System_3d.cpp:633-661); also hands the slice
// counter to the next step.
__global__ void k_mx_out(double *ex, double *ey, double *ez, const double *psi, const double *psi_prev,
                         const double *by, const double *bz, int n1, int n2, double dx, double dy, double dz,
                         int with_ex, Ctl *ctl, int slot, int final_step) {
    if (!final_step) {
        MX_SKIP(ctl);
    } else if (ctl->stop_acc) {
        by = nullptr;
        bz = nullptr;
    }
    if (!final_step && blockIdx.x == 0 && threadIdx.x == 0) ctl->slice_of[1 - slot] = ctl->slice_of[slot] - 1;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n1 * n2) return;
    const int j = (int)(t / n2), k = (int)(t % n2);
    if (with_ex) ex[t] = (psi[t] - psi_prev[t]) / dx;
    const size_t tj = j < n1 - 1 ? t + n2 : (size_t)k;
    const size_t tk = k < n2 - 1 ? t + 1 : (size_t)n2 * j;
    ey[t] = -(psi[tj] - psi[t]) / dy + (bz ? bz[t] : 0.0);
    ez[t] = -(psi[tk] - psi[t]) / dz - (by ? by[t] : 0.0);
}

// ------------------------------------------------------- solver functors
// psi source: jx_bunch(i) - rho_bunch(i) - rho_ion + deposits; clears the accumulator.
struct SrcPsi {
    double *acc;
    const double *rho_ion, *rho_b3, *jx_b3;
    const Ctl *ctl;
    int slot, n2;
    size_t nn;
    __device__ double operator()(int j, int k) const {
        const size_t t = (size_t)j * n2 + k;
        const int slice = ctl->slice_of[slot];
        const double jb = jx_b3 ? jx_b3[(size_t)slice * nn + t] : 0.0;
        const double rb = rho_b3 ? rho_b3[(size_t)slice * nn + t] : 0.0;
        const double v = acc[t];
        acc[t] = 0.0;
        return (jb - rb - rho_ion[t]) + v;
    }
};

// B_y (blockIdx.y = 0, :530-536) and B_z (blockIdx.y = 1, :555-564) sources.
struct SrcB {
    const double *by, *bz, *jx, *jy, *jz, *jyn, *jzn;
    int n1, n2;
    double D, dx, dy, dz;
    __device__ double operator()(int j, int k) const {
        const size_t t = (size_t)j * n2 + k;
        if (blockIdx.y == 0) {
            const double jn = k < n2 - 1 ? jx[t + 1] : jx[(size_t)j * n2];
            const double djz = (jz[t] - jzn[t]) / dx;
            return -D * by[t] - (jn - jx[t]) / dz + djz;
        }
        const double djy = (jy[t] - jyn[t]) / dx;
        if (j < n1 - 1) return -D * bz[t] + (jx[t + n2] - jx[t]) / dy - djy;
        return -D * bz[k] + (jx[k] - jx[t]) / dy - djy;
    }
};

// psi finalisation: v/n1/n2, minus the border mean (normalize_potential,
// :30-47) computed from the column pass's partial sums, then the clamp.
struct SinkPsiFinal {
    double *dst;
    const double *bpart;  // [ncb] (P1, P2) pairs then G00
    int ncb, n1, n2;
    double thr;
    double b;
    __device__ void begin() {
        double s1 = 0.0, s2 = 0.0;
        for (int i = threadIdx.x; i < ncb; i += blockDim.x) {
            s1 += bpart[2 * i];
            s2 += bpart[2 * i + 1];
        }
        s1 = block_sum(s1);
        s2 = block_sum(s2);
        const double g00 = bpart[2 * ncb];
        const double sraw = s1 + (double)n2 * g00 - s2;
        b = sraw / n1 / n2 / (n1 + n2 - 1);
    }
    __device__ void operator()(int j, int k, double v) const {
        double r = v / n1 / n2 - b;
        if (r < thr) r = thr;
        dst[(size_t)j * n2 + k] = r;
    }
};

// B store + stop rule, paired: blockIdx.y selects B_y / B_z.
struct SinkBPair {
    double *by, *bz;
    int n1, n2;
    double thr;
    Ctl *ctl;
    int slot;
    __device__ void begin() {}
    __device__ void operator()(int j, int k, double v) const {
        const double b = v / n1 / n2;
        (blockIdx.y == 0 ? by : bz)[(size_t)j * n2 + k] = b;
        if (fabs(b) > thr || isnan(b)) {
            ctl->stop_acc = 1;
            ctl->stop_slice = ctl->slice_of[slot];
        }
    }
};

// Column pass of the psi solves, with the partial sums the border mean needs:
// after the inverse column transform G[j][l], the c2r row value at k = 0 is
// sum_l c_l Re G[j][l] (c_0 = c_{n2/2} = 1, else 2) and sum_k x[0][k] = n2 Re G[0][0].
__global__ __launch_bounds__(kSolverThreads) void k_cols_psi(double2 *spec, int n1, int nh, int n2, FftAxisDev ax,
                                                             const double *lam1, const double *lam2, int C,
                                                             double *bpart, const Ctl *ctl) {
    if (solver_skip(ctl, 1)) return;
    extern __shared__ double2 lds2[];
    double2 *a = lds2, *b = lds2 + (size_t)C * n1, *twl = lds2 + (size_t)2 * C * n1;
    double *lam1l = reinterpret_cast<double *>(twl + n1);
    const int l0 = blockIdx.x * C;
    lds_load_tw(twl, ax);
    for (int t = threadIdx.x; t < n1; t += blockDim.x) lam1l[t] = lam1[t];
    for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
        const int c = t % C, j = t / C, l = l0 + c;
        a[c * n1 + j] = l < nh ? spec[(size_t)j * nh + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    double2 *r = lds_fft(a, b, C, ax, twl, -1);
    for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
        const int c = t / n1, m = t - c * n1, l = l0 + c;
        if (l >= nh) continue;
        double2 v = r[t];
        if (m == 0 && l == 0) {
            v = make_double2(0.0, 0.0);
        } else {
            const double mult = -0.0 + lam1l[m] + lam2[l];
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
    // partial sums over the whole block: sum_c w_l sum_j Re G[j][l] and sum_c w_l Re G[0][l]
    double s1 = 0.0, s2 = 0.0;
    for (int t = threadIdx.x; t < C * n1; t += blockDim.x) {
        const int c = t / n1, j = t - c * n1, l = l0 + c;
        if (l >= nh) continue;
        const double wl = (l == 0 || 2 * l == n2) ? 1.0 : 2.0;
        const double v = wl * o[t].x;
        s1 += v;
        if (j == 0) s2 += v;
    }
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    if (threadIdx.x == 0) {
        bpart[2 * blockIdx.x] = s1;
        bpart[2 * blockIdx.x + 1] = s2;
        if (blockIdx.x == 0) bpart[2 * gridDim.x] = o[0].x;
    }
}

// ------------------------------------------------------------ re-binning
// Counting sort of the particles by 16x16-cell tile (stable order is not
// needed: the deposits are order-independent up to rounding).
constexpr int kTile = 16;
__global__ void k_sort_hist(MxPart P, unsigned *count, int ntk) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    const uint32_t c = P.cell[p];
    const int t = (int)((c >> 16) / kTile) * ntk + (int)((c & 0xffffu) / kTile);
    atomicAdd(&count[t], 1u);
}

__global__ void k_sort_scan(unsigned *count, unsigned *cursor, int ntiles) {
    // single block exclusive scan; count becomes the offsets
    __shared__ unsigned part[1024];
    const int per = (ntiles + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per;
    unsigned s = 0;
    for (int i = b; i < b + per && i < ntiles; ++i) s += count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            const unsigned v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    unsigned run = part[threadIdx.x];
    for (int i = b; i < b + per && i < ntiles; ++i) {
        const unsigned v = count[i];
        count[i] = run;
        cursor[i] = 0;
        run += v;
    }
}

__global__ void k_sort_scatter(MxPart P, MxPart Q, const unsigned *offset, unsigned *cursor, const int *perm,
                               int *perm2, int ntk) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    const uint32_t c = P.cell[p];
    const int t = (int)((c >> 16) / kTile) * ntk + (int)((c & 0xffffu) / kTile);
    const unsigned d = offset[t] + atomicAdd(&cursor[t], 1u);
    if (d >= (unsigned)P.np) return;
    Q.cell[d] = c;
    Q.oy[d] = P.oy[p];
    Q.oz[d] = P.oz[p];
    Q.py[d] = P.py[p];
    Q.pz[d] = P.pz[p];
    Q.w[d] = P.w[p];
    Q.dpy[d] = P.dpy[p];
    Q.dpz[d] = P.dpz[p];
    perm2[d] = perm[p];
}

__global__ void k_sort_copyback(MxPart P, MxPart Q, int *perm, const int *perm2) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    P.cell[p] = Q.cell[p];
    P.oy[p] = Q.oy[p];
    P.oz[p] = Q.oz[p];
    P.py[p] = Q.py[p];
    P.pz[p] = Q.pz[p];
    P.w[p] = Q.w[p];
    P.dpy[p] = Q.dpy[p];
    P.dpz[p] = Q.dpz[p];
    perm[p] = perm2[p];
}

// ------------------------------------------------------------------ host
struct MixedState {
    MxPart P{}, Q{}, I{};   // live, re-binning scratch, initial state of a pass
    int *perm = nullptr, *perm2 = nullptr, *perm0 = nullptr;
    bool init_valid = false;
    unsigned *count = nullptr, *cursor = nullptr;
    double *bpart = nullptr;
    double *jbuf[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [parity][y/z]
    double *ex = nullptr;
    std::vector<int> h_perm;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int ncb = 0;
    size_t cap = 0;
};

static MixedState *mstate(Wake *w) { return static_cast<MixedState *>(w->mixed); }

static void free_part(MxPart &P) {
    void *arrs[] = {P.cell, P.oy, P.oz, P.py, P.pz, P.w, P.dpy, P.dpz};
    for (void *a : arrs)
        if (a) (void)hipFree(a);
    P = MxPart{};
}

static int alloc_part(MxPart &P, size_t n) {
    n = n ? n : 1;
    QP_HIP_CHECK(hipMalloc(&P.cell, 4 * n));
    QP_HIP_CHECK(hipMalloc(&P.oy, 4 * n));
    QP_HIP_CHECK(hipMalloc(&P.oz, 4 * n));
    QP_HIP_CHECK(hipMalloc(&P.py, 8 * n));
    QP_HIP_CHECK(hipMalloc(&P.pz, 8 * n));
    QP_HIP_CHECK(hipMalloc(&P.w, 4 * n));
    QP_HIP_CHECK(hipMalloc(&P.dpy, 4 * n));
    QP_HIP_CHECK(hipMalloc(&P.dpz, 4 * n));
    return QP_OK;
}

void wake_destroy_mixed(Wake *w) {
    MixedState *m = mstate(w);
    if (!m) return;
    free_part(m->P);
    free_part(m->Q);
    free_part(m->I);
    void *arrs[] = {m->perm, m->perm2, m->perm0, m->count, m->cursor, m->bpart, m->jbuf[0][0], m->jbuf[0][1],
                    m->jbuf[1][0], m->jbuf[1][1]};
    for (void *a : arrs)
        if (a) (void)hipFree(a);
    for (auto &g : m->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    delete m;
    w->mixed = nullptr;
}

static int ensure_state(Wake *w) {
    MixedState *m = mstate(w);
    const size_t nn = w->nn();
    if (!m) {
        m = new MixedState();
        w->mixed = m;
        const int ncb = w->S.wave ? wcols_blocks(w->S.n1, w->S.nh)
                                  : (w->S.nh + w->S.cols_per_block - 1) / w->S.cols_per_block;
        if (w->S.cols_per_block > 64) {
            set_error("column block too wide for the border partial sums");
            return QP_ERR_STATE;
        }
        m->ncb = ncb;
        QP_HIP_CHECK(hipMalloc(&m->bpart, sizeof(double) * (2 * ncb + 2)));
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) QP_HIP_CHECK(hipMalloc(&m->jbuf[a][b], sizeof(double) * nn));
        const int ntk = (w->p.nz + kTile - 1) / kTile, ntj = (w->p.ny + kTile - 1) / kTile;
        QP_HIP_CHECK(hipMalloc(&m->count, sizeof(unsigned) * ntk * ntj));
        QP_HIP_CHECK(hipMalloc(&m->cursor, sizeof(unsigned) * ntk * ntj));
    }
    if (m->cap < w->np || !m->P.cell) {
        free_part(m->P);
        free_part(m->Q);
        free_part(m->I);
        if (m->perm) (void)hipFree(m->perm);
        if (m->perm2) (void)hipFree(m->perm2);
        if (m->perm0) (void)hipFree(m->perm0);
        QP_TRY(alloc_part(m->P, w->np));
        QP_TRY(alloc_part(m->Q, w->np));
        QP_TRY(alloc_part(m->I, w->np));
        QP_HIP_CHECK(hipMalloc(&m->perm, sizeof(int) * (w->np ? w->np : 1)));
        QP_HIP_CHECK(hipMalloc(&m->perm2, sizeof(int) * (w->np ? w->np : 1)));
        QP_HIP_CHECK(hipMalloc(&m->perm0, sizeof(int) * (w->np ? w->np : 1)));
        m->cap = w->np;
        m->init_valid = false;
        for (auto &g : m->gexec)
            if (g) (void)hipGraphExecDestroy(g);
        m->gexec[0] = m->gexec[1] = nullptr;
    }
    m->P.np = m->Q.np = m->I.np = (int)w->np;
    return QP_OK;
}

static inline void to_cells(double y, double d, int n, int &j, uint32_t &o) {
    double Y = y / d;
    double fl = std::floor(Y);
    int jj = (int)fl;
    double u = std::rint((Y - fl) * kTwo32);
    if (u >= kTwo32) {
        u -= kTwo32;
        jj += 1;
    }
    if (jj >= n) jj -= n;
    if (jj < 0 || jj >= n) jj = 0;
    j = jj;
    o = (uint32_t)u;
}

// Convert the host FP64 particles to the device layout once (sorted by
// 16x16-cell tile, row-major inside a tile) into the pass's initial-state
// buffers; every pass then starts with device-to-device copies, as the
// reference re-initialises its particles each pass (src/System_3d.cpp:262).
static int build_initial_state(Wake *w, MixedState *m) {
    const size_t np = w->np;
    const int n1 = w->p.ny, n2 = w->p.nz;
    const int ntk = (n2 + kTile - 1) / kTile;
    std::vector<uint32_t> cell(np), oy(np), oz(np);
    std::vector<long long> key(np);
    for (size_t i = 0; i < np; ++i) {
        int j, k;
        to_cells(w->h_y[i], w->p.dy, n1, j, oy[i]);
        to_cells(w->h_z[i], w->p.dz, n2, k, oz[i]);
        cell[i] = ((uint32_t)j << 16) | (uint32_t)k;
        const long long tile = (long long)(j / kTile) * ntk + (k / kTile);
        key[i] = (tile * n1 + j) * (long long)n2 + k;
    }
    m->h_perm.resize(np);
    std::iota(m->h_perm.begin(), m->h_perm.end(), 0);
    std::stable_sort(m->h_perm.begin(), m->h_perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<uint32_t> c2(np), y2(np), z2(np);
    std::vector<double> py(np), pz(np);
    std::vector<float> wt(np);
    for (size_t s = 0; s < np; ++s) {
        const int i = m->h_perm[s];
        c2[s] = cell[i];
        y2[s] = oy[i];
        z2[s] = oz[i];
        py[s] = w->h_py[i];
        pz[s] = w->h_pz[i];
        wt[s] = (float)w->h_w[i];
    }
    QP_HIP_CHECK(hipMemcpy(m->I.cell, c2.data(), 4 * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(m->I.oy, y2.data(), 4 * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(m->I.oz, z2.data(), 4 * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(m->I.py, py.data(), 8 * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(m->I.pz, pz.data(), 8 * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(m->I.w, wt.data(), 4 * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(m->perm0, m->h_perm.data(), sizeof(int) * np, hipMemcpyHostToDevice));
    m->init_valid = true;
    return QP_OK;
}

int wake_upload_mixed(Wake *w) {
    QP_TRY(ensure_state(w));
    MixedState *m = mstate(w);
    if (w->particles_dirty || !m->init_valid) {
        QP_TRY(build_initial_state(w, m));
        w->particles_dirty = false;
    }
    const size_t np = w->np;
    hipStream_t s = w->stream();
    QP_HIP_CHECK(hipMemcpyAsync(m->P.cell, m->I.cell, 4 * np, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.oy, m->I.oy, 4 * np, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.oz, m->I.oz, 4 * np, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.py, m->I.py, 8 * np, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.pz, m->I.pz, 8 * np, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.w, m->I.w, 4 * np, hipMemcpyDeviceToDevice, s));
    // py_next = pz_next = 0 (src/System_3d.cpp:740-741): the first slice reads no
    // increment, and with K = 0 the commit writes p = 0 directly
    QP_HIP_CHECK(hipMemsetAsync(m->P.dpy, 0, 4 * np, s));
    QP_HIP_CHECK(hipMemsetAsync(m->P.dpz, 0, 4 * np, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->perm, m->perm0, sizeof(int) * np, hipMemcpyDeviceToDevice, s));
    return QP_OK;
}

int wake_download_mixed(Wake *w, double *y, double *z, double *py, double *pz) {
    MixedState *m = mstate(w);
    if (!m) {
        set_error("mixed state not initialised");
        return QP_ERR_STATE;
    }
    const size_t np = w->np;
    QP_HIP_CHECK(hipStreamSynchronize(w->stream()));
    std::vector<uint32_t> c(np), oy(np), oz(np);
    std::vector<double> fy(np), fz(np);
    std::vector<int> perm(np);
    QP_HIP_CHECK(hipMemcpy(c.data(), m->P.cell, 4 * np, hipMemcpyDeviceToHost));
    QP_HIP_CHECK(hipMemcpy(oy.data(), m->P.oy, 4 * np, hipMemcpyDeviceToHost));
    QP_HIP_CHECK(hipMemcpy(oz.data(), m->P.oz, 4 * np, hipMemcpyDeviceToHost));
    QP_HIP_CHECK(hipMemcpy(fy.data(), m->P.py, 8 * np, hipMemcpyDeviceToHost));
    QP_HIP_CHECK(hipMemcpy(fz.data(), m->P.pz, 8 * np, hipMemcpyDeviceToHost));
    QP_HIP_CHECK(hipMemcpy(perm.data(), m->perm, sizeof(int) * np, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < np; ++s) {
        const int i = perm[s];
        if (i < 0 || (size_t)i >= np) continue;
        if (y) y[i] = ((double)(c[s] >> 16) + oy[s] * kInvTwo32) * w->p.dy;
        if (z) z[i] = ((double)(c[s] & 0xffffu) + oz[s] * kInvTwo32) * w->p.dz;
        if (py) py[i] = fy[s];
        if (pz) pz[i] = fz[s];
    }
    return QP_OK;
}

namespace {

struct MxRun {
    Wake *w;
    MixedState *m;
    MxGeom g;
    MxSliceIn si[2];
    size_t nn;
    int np, pb, gb;
    size_t ldsC, ldsS;
};

#define MLAUNCH(kern, grid, block, lds, ...)                                            \
    do {                                                                                \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, s, __VA_ARGS__);         \
        QP_HIP_CHECK(hipGetLastError());                                                \
    } while (0)

int psi_solve(MxRun &R, hipStream_t s, int slot, double *dst) {
    Wake *w = R.w;
    const Solver2d &S = w->S;
    MixedState *m = R.m;
    SrcPsi src{w->src, w->rho_ion, w->rho_b3, w->jx_b3, w->ctl, slot, S.n2, R.nn};
    SinkPsiFinal snk{dst, m->bpart, m->ncb, S.n1, S.n2, w->p.psi_threshold - 1, 0.0};
    if (S.wave) {
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_fwd(S, src, S.spec[0], S.spec[0], 1, s, (const Ctl *)w->ctl, 1));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_cols(S, 0.0, 1, m->bpart, s, (const Ctl *)w->ctl, 1));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_inv(S, (const double2 *)S.spec[0], (const double2 *)S.spec[0], snk, 1, s,
                                (const Ctl *)w->ctl, 1));
        prof_end(w, "solve");
        return QP_OK;
    }
    prof_begin(w, "solve");
    MLAUNCH(k_rows_fwd<SrcPsi>, rows_blocks(S), kSolverThreads, rows_lds(S), src, S.spec[0], S.spec[0], S.n1, S.n2,
            S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)w->ctl, 1);
    prof_end(w, "solve");
    prof_begin(w, "solve");
    MLAUNCH(k_cols_psi, m->ncb, kSolverThreads, cols_lds(S), S.spec[0], S.n1, S.nh, S.n2, S.ax1.dev,
            (const double *)S.d_lam1, (const double *)S.d_lam2, S.cols_per_block, m->bpart, (const Ctl *)w->ctl);
    prof_end(w, "solve");
    prof_begin(w, "solve");
    MLAUNCH(k_rows_inv<SinkPsiFinal>, rows_blocks(S), kSolverThreads, rows_lds(S), (const double2 *)S.spec[0],
            (const double2 *)S.spec[0], snk, S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)w->ctl, 1);
    prof_end(w, "solve");
    return QP_OK;
}

int b_solve(MxRun &R, hipStream_t s, int slot) {
    Wake *w = R.w;
    const Solver2d &S = w->S;
    MixedState *m = R.m;
    const qp_wake_params &p = w->p;
    double *jy = m->jbuf[1 - slot][0], *jz = m->jbuf[1 - slot][1];
    double *jyn = m->jbuf[slot][0], *jzn = m->jbuf[slot][1];
    SrcB src{w->by, w->bz, w->jx, jy, jz, jyn, jzn, S.n1, S.n2, p.D, p.dx, p.dy, p.dz};
    SinkBPair snk{w->by, w->bz, S.n1, S.n2, p.b_threshold, w->ctl, slot};
    if (S.wave) {
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_fwd(S, src, S.spec[0], S.spec[1], 2, s, (const Ctl *)w->ctl, 2));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_cols(S, p.D, 2, nullptr, s, (const Ctl *)w->ctl, 2));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_inv(S, (const double2 *)S.spec[0], (const double2 *)S.spec[1], snk, 2, s,
                                (const Ctl *)w->ctl, 2));
        prof_end(w, "solve");
        return QP_OK;
    }
    prof_begin(w, "solve");
    hipLaunchKernelGGL(k_rows_fwd<SrcB>, dim3(rows_blocks(S), 2), dim3(kSolverThreads), rows_lds(S), s, src,
                       S.spec[0], S.spec[1], S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)w->ctl, 2);
    QP_HIP_CHECK(hipGetLastError());
    prof_end(w, "solve");
    prof_begin(w, "solve");
    QP_TRY(solve_cols(S, p.D, 2, s, w->ctl, 2));
    prof_end(w, "solve");
    prof_begin(w, "solve");
    hipLaunchKernelGGL(k_rows_inv<SinkBPair>, dim3(rows_blocks(S), 2), dim3(kSolverThreads), rows_lds(S), s,
                       (const double2 *)S.spec[0], (const double2 *)S.spec[1], snk, S.n1, S.n2, S.nh, S.ax2.dev,
                       S.rows_per_block, (const Ctl *)w->ctl, 2);
    QP_HIP_CHECK(hipGetLastError());
    prof_end(w, "solve");
    return QP_OK;
}

// Enqueue the work of slice i (everything between two output steps).
int enqueue_slice(MxRun &R, hipStream_t s, int i) {
    Wake *w = R.w;
    MixedState *m = R.m;
    const qp_wake_params &p = w->p;
    const int slot = i & 1;
    const bool first = (i == p.nx - 1);
    const bool s_phase = (i == 0) || (p.K == 0);
    const MxSliceIn &si = R.si[slot];
    double *psi = w->psi_b[slot];
    const int laser = (w->laser_set && !w->a_sqr3) ? 1 : 0;
    prof_begin(w, "push_A");
    MLAUNCH(k_mx_A, R.pb, kPT, 0, m->P, R.g, si, (const double *)w->psim, w->src, w->asq2, w->laser, laser, w->rho,
            w->jx, w->chi, s_phase ? 1 : 0, first ? 0 : (p.K == 0 ? 2 : 1), w->ctl);
    prof_end(w, "push_A");
    QP_TRY(psi_solve(R, s, slot, psi));
    if (s_phase) {
        prof_begin(w, "push_S");
        MLAUNCH(k_mx_S, R.pb, kPT, R.ldsS, m->P, R.g, si, (const double *)psi, w->rho, w->jx, w->chi, w->ctl);
        prof_end(w, "push_S");
        if (i == 0) return QP_OK;
    } else {
        for (int it = 0; it < p.K; ++it) {
            prof_begin(w, "push_B");
            MLAUNCH(k_mx_B, R.pb, kPT, 0, m->P, R.g, si, (const double *)psi, (const double *)w->by,
                    (const double *)w->bz, w->src, w->rho, w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1],
                    it == 0 ? 1 : 0, w->ctl);
            prof_end(w, "push_B");
            QP_TRY(psi_solve(R, s, slot, w->psim));
            prof_begin(w, "push_C");
            MLAUNCH(k_mx_C, R.pb, kPT, R.ldsC, m->P, R.g, si, (const double *)psi, (const double *)w->psim, w->rho,
                    w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1], w->ctl);
            prof_end(w, "push_C");
            QP_TRY(b_solve(R, s, slot));
        }
    }
    MLAUNCH(k_mx_out, R.gb, 256, 0, w->ex, w->ey, w->ez, (const double *)psi, (const double *)w->psi_b[1 - slot],
            (const double *)w->by, (const double *)w->bz, p.ny, p.nz, p.dx, p.dy, p.dz, first ? 0 : 1, w->ctl, slot,
            0);
    return QP_OK;
}

int resort(MxRun &R, hipStream_t s) {
    Wake *w = R.w;
    MixedState *m = R.m;
    const int ntk = (w->p.nz + kTile - 1) / kTile, ntj = (w->p.ny + kTile - 1) / kTile;
    const int ntiles = ntk * ntj;
    QP_HIP_CHECK(hipMemsetAsync(m->count, 0, sizeof(unsigned) * ntiles, s));
    MLAUNCH(k_sort_hist, R.pb, kPT, 0, m->P, m->count, ntk);
    MLAUNCH(k_sort_scan, 1, 1024, 0, m->count, m->cursor, ntiles);
    MLAUNCH(k_sort_scatter, R.pb, kPT, 0, m->P, m->Q, (const unsigned *)m->count, m->cursor, (const int *)m->perm,
            m->perm2, ntk);
    MLAUNCH(k_sort_copyback, R.pb, kPT, 0, m->P, m->Q, m->perm, (const int *)m->perm2);
    return QP_OK;
}

}  // namespace

int wake_run_mixed(Wake *w, qp_slice_cb cb, void *user, int *stop_slice) {
    const qp_wake_params &p = w->p;
    hipStream_t s = w->stream();
    QP_TRY(wake_upload_mixed(w));
    MixedState *m = mstate(w);
    MxRun R;
    R.w = w;
    R.m = m;
    R.nn = w->nn();
    R.np = (int)w->np;
    R.pb = pblocks(R.np);
    R.gb = (int)((R.nn + 255) / 256);
    R.g = MxGeom{p.ny, p.nz, p.dy, p.dz, p.ly, p.lz, p.dx, 1.0 / p.dy, 1.0 / p.dz};
    for (int sl = 0; sl < 2; ++sl) R.si[sl] = MxSliceIn{w->a_sqr3, w->asq2, w->rho_b3, w->jx_b3, sl};
    R.ldsC = sizeof(double) * 5 * kWinCap;
    R.ldsS = sizeof(double) * 3 * kWinCap;

    // zero the fields (src/System_3d.cpp:265-316) and the control block
    double *zero_list[] = {w->psi_b[0], w->psi_b[1], w->psim, w->by, w->bz, w->rho_ion, w->src, w->rho, w->jx,
                           w->chi, w->ex, w->ey, w->ez, w->asq2};
    for (double *a : zero_list) QP_HIP_CHECK(hipMemsetAsync(a, 0, sizeof(double) * R.nn, s));
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) QP_HIP_CHECK(hipMemsetAsync(m->jbuf[a][b], 0, sizeof(double) * R.nn, s));
    Ctl hc0{};
    hc0.slice_of[(p.nx - 1) & 1] = p.nx - 1;
    hc0.slice_of[p.nx & 1] = p.nx;
    QP_HIP_CHECK(hipMemcpyAsync(w->ctl, &hc0, sizeof(Ctl), hipMemcpyHostToDevice, s));
    QP_HIP_CHECK(hipEventRecord(w->ev0, s));
    if (R.np) MLAUNCH(k_mx_ion, R.pb, kPT, 0, m->P, R.g, w->rho_ion);

    const bool graph_ok = p.use_graph && !cb && !w->prof_class;
    int delivered_stop = 0;
    for (int i = p.nx - 1; i >= 0; --i) {
        const int slot = i & 1;
        if (p.resort_every > 0 && i < p.nx - 1 && ((p.nx - 1 - i) % p.resort_every) == 0 && R.np)
            QP_TRY(resort(R, s));
        const bool generic = (i < p.nx - 1) && (i > 0) && (p.K > 0);
        if (graph_ok && generic) {
            if (!m->gexec[slot]) {
                hipGraph_t gr;
                QP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int st = enqueue_slice(R, s, i);
                hipError_t e = hipStreamEndCapture(s, &gr);
                if (st < 0) return st;
                if (e != hipSuccess) {
                    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                    return QP_ERR_HIP;
                }
                e = hipGraphInstantiate(&m->gexec[slot], gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) {
                    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                    return QP_ERR_HIP;
                }
            }
            QP_HIP_CHECK(hipGraphLaunch(m->gexec[slot], s));
        } else {
            QP_TRY(enqueue_slice(R, s, i));
        }
        if (cb && i > 0) {
            QP_HIP_CHECK(hipStreamSynchronize(s));
            Ctl hc;
            QP_HIP_CHECK(hipMemcpy(&hc, w->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
            if (hc.stop_acc) {
                QP_TRY(deliver_slice(w, cb, user, hc.stop_slice, QP_SLICE_STOPPED, nullptr, nullptr, nullptr,
                                     nullptr, nullptr, nullptr, nullptr, nullptr, w->by, w->bz, w->chi));
                delivered_stop = 1;
                break;
            }
            QP_TRY(deliver_slice(w, cb, user, i, QP_SLICE_NORMAL, w->psi_b[slot], w->rho, w->jx,
                                 m->jbuf[1 - slot][0], m->jbuf[1 - slot][1], w->ex, w->ey, w->ez, w->by, w->bz,
                                 w->chi));
        }
    }
    QP_HIP_CHECK(hipStreamSynchronize(s));
    Ctl hc;
    QP_HIP_CHECK(hipMemcpy(&hc, w->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    // output_step(0) after the loop (src/System_3d.cpp:622)
    const int cur = hc.stop_acc ? hc.stop_slice : 0;
    const int cslot = cur & 1;
    hipLaunchKernelGGL(k_mx_out, dim3(R.gb), dim3(256), 0, s, w->ex, w->ey, w->ez, (const double *)w->psi_b[cslot],
                       (const double *)w->psi_b[1 - cslot], (const double *)w->by, (const double *)w->bz, p.ny, p.nz,
                       p.dx, p.dy, p.dz, p.nx > 1 ? 1 : 0, w->ctl, 0, 1);
    QP_HIP_CHECK(hipGetLastError());
    QP_HIP_CHECK(hipEventRecord(w->ev1, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    QP_HIP_CHECK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    w->last_ms = ms;
    if (hc.stop_acc && cb && !delivered_stop) {
        QP_TRY(deliver_slice(w, cb, user, hc.stop_slice, QP_SLICE_STOPPED, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr, w->by, w->bz, w->chi));
    }
    if (cb) {
        const double *by0 = hc.stop_acc ? w->zero2 : w->by;
        const double *bz0 = hc.stop_acc ? w->zero2 : w->bz;
        const double *chi0 = hc.stop_acc ? w->zero2 : w->chi;
        QP_TRY(deliver_slice(w, cb, user, 0, QP_SLICE_FINAL, w->psi_b[cslot], w->rho, w->jx, m->jbuf[1 - cslot][0],
                             m->jbuf[1 - cslot][1], w->ex, w->ey, w->ez, by0, bz0, chi0));
    }
    if (stop_slice) *stop_slice = hc.stop_acc ? hc.stop_slice : -1;
    return hc.stop_acc ? QP_STOPPED : QP_OK;
}

}  // namespace qp
