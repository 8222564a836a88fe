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
#include <atomic>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "qp_mixed_kernels.h"
#include "qp_wave_solver.h"

namespace qp {


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
// Batches: blockIdx.y is the member; its grids sit at + member * nn, its
// particles at + member * pstride, its control block at ctl + member.
template <int PT, int CAP>
__device__ __forceinline__ void mx_A(MxPart P, MxGeom g, MxSliceIn si, const double *psim, double *src_acc,
                                              double *asq2, const LaserGauss *Ls, int laser, double *rho, double *jx,
                                              double *chi, int init_rho, int advance, Ctl *ctl, int bx, int wk, int gx) {
    // The particle's state is loaded before the control block is checked: the
    // index stays inside the member's allocation (pstride is a whole number of
    // blocks), so these loads share a round trip with the control block's.
    const int lp = bx * PT + threadIdx.x;
    const unsigned p = (unsigned)(wk * P.pstride + lp);
    const Pos q0 = load_pos(P, p);
    const double w0 = P.w[p];
    const double2 pst = P.ps[p];
    ctl += wk;
    MX_SKIP(ctl);
    __shared__ double lds[CAP];
    __shared__ int sh[4 * (PT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2, go = (size_t)wk * nn;
    psim += go;
    src_acc += go;
    asq2 += go;
    rho += go;
    jx += go;
    chi += go;
    const LaserGauss L = Ls[wk];
    const int slice = mx_slice(ctl, si);
    const bool valid = lp < mx_np(P, wk);
    Pos q{0, 0, 0.0, 0.0};
    double w = 0.0;
    if (valid) {
        q = q0;
        w = w0;
        if (advance) {
            // commit p <- p* (advance 2: K = 0, where p* keeps its initial value 0)
            const double pyv = advance != 2 ? pst.x : 0.0, pzv = advance != 2 ? pst.y : 0.0;
            const double ps = mx_gather(psim, q, n1, n2);
            P.p[p] = make_double2(pyv, pzv);
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
            P.off[p] = make_uint2(oy, oz);
            q.j = j;
            q.k = k;
            q.fy = (double)oy * kInvTwo32;
            q.fz = (double)oz * kInvTwo32;
        }
    }
    if (threadIdx.x == 0) {
        sh[4 * (PT / 64)] = q.j;
        sh[4 * (PT / 64) + 1] = q.k;
    }
    __syncthreads();
    const int jref = sh[4 * (PT / 64)], kref = sh[4 * (PT / 64) + 1];
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) range_add(q.j, q.k, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    const Win win = block_window<CAP>(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_note_overflow(win, ctl);
    win_zero<CAP>(lds, 1, win);
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
    win_flush<CAP>(lds, gl, 1, win, n1, n2);
    // The slice's laser plane and the S-phase initial rho / jx / chi last, so that
    // these stores do not queue ahead of the particle loads (nothing in this kernel
    // reads them).
    if (laser || init_rho) {
        const double *rb = si.rho_b3 ? si.rho_b3 + (size_t)slice * si.bstride : nullptr;
        const double *jb = si.jx_b3 ? si.jx_b3 + (size_t)slice * si.bstride : nullptr;
        for (size_t t = (size_t)bx * PT + threadIdx.x; t < nn; t += (size_t)gx * PT) {
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
}
template <int PT, int CAP>
__global__ __launch_bounds__(PT) void k_mx_A(MxPart P, MxGeom g, MxSliceIn si, const double *psim, double *src_acc,
                                              double *asq2, const LaserGauss *Ls, int laser, double *rho, double *jx,
                                              double *chi, int init_rho, int advance, Ctl *ctl) {
    mx_A<PT, CAP>(P, g, si, psim, src_acc, asq2, Ls, laser, rho, jx, chi, init_rho, advance, ctl, blockIdx.x, blockIdx.y,
         gridDim.x);
}

// ------------------------------------------------------------------ B phase
// gamma = (1 + p^2 + a + (1+psi)^2) / (2 (1+psi)) (src/System_3d.cpp:373-374), with 1/(1+psi) given.
__device__ __forceinline__ double gamma_of(double p2, double a, double ps, double inv1ps) {
    return 0.5 * (1 + p2 + a + (1 + ps) * (1 + ps)) * inv1ps;
}

template <int PT, int CAP>
__device__ __forceinline__ void mx_B(MxPart P, MxGeom g, MxSliceIn si, const double *psi, const double *by,
                                              const double *bz, double *src_acc, double *rho, double *jx,
                                              double *chi, double *jyn, double *jzn, int first_iter, int init_rho,
                                              Ctl *ctl, int bx, int wk, int gx) {
    ctl += wk;
    MX_SKIP(ctl);
    __shared__ double lds[CAP];
    __shared__ int sh[4 * (PT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2, go = (size_t)wk * nn;
    psi += go;
    by += go;
    bz += go;
    src_acc += go;
    rho += go;
    jx += go;
    chi += go;
    jyn += go;
    jzn += go;
    const int slice = mx_slice(ctl, si);
    const double *asq = mx_asq(ctl, si, nn, go);
    const int lp = bx * PT + threadIdx.x;
    const bool valid = lp < mx_np(P, wk);
    const unsigned p = (unsigned)(wk * P.pstride + lp);
    Pos q{0, 0, 0.0, 0.0}, qm{0, 0, 0.0, 0.0};
    double w = 0.0;
    double2 pp = make_double2(0.0, 0.0), pst = pp;
    if (valid) {
        q = load_pos(P, p);
        w = P.w[p];
        pp = P.p[p];
        pst = P.ps[p];  // P6 of the previous iteration (:503-505); unused in the first
    }
    // Gather window: rows and columns -1 .. +2 around the block's cells hold every
    // node of the value + gradient stencils and of the staggered B stencils.
    __shared__ double gw[4 * kGwCap];
    Win gwn;
    {
        if (threadIdx.x == 0) {
            sh[4 * (PT / 64)] = q.j;
            sh[4 * (PT / 64) + 1] = q.k;
        }
        __syncthreads();
        const int jref = sh[4 * (PT / 64)], kref = sh[4 * (PT / 64) + 1];
        int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
        if (valid) {
            const int r = rel_idx(q.j, jref, n1), c = rel_idx(q.k, kref, n2);
            rmin = r - 1;
            rmax = r + 2;
            cmin = c - 1;
            cmax = c + 2;
        }
        gwn = block_window<kGwCap>(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
        if (gwn.lds) {
            const double *src[4] = {psi, asq, by, bz};
            gwin_load<4>(gw, kGwCap, src, gwn, n1, n2);
        }
        __syncthreads();
    }
    if (valid) {
        ValGrad ga, gp;
        double byp, bzp;
        if (gwn.lds) {
            int r, c;
            gwin_rc(gwn, q.j, q.k, n1, n2, r, c);
            gp = lg_vg(gw, gwn.wk, r, c, q.fy, q.fz, g.inv_d1, g.inv_d2);
            ga = lg_vg(gw + kGwCap, gwn.wk, r, c, q.fy, q.fz, g.inv_d1, g.inv_d2);
            byp = lg_hz(gw + 2 * kGwCap, gwn.wk, r, c, q.fy, q.fz);
            bzp = lg_hy(gw + 3 * kGwCap, gwn.wk, r, c, q.fy, q.fz);
        } else {
            ga = mx_gather_vg(asq, q, n1, n2, g.inv_d1, g.inv_d2);
            gp = mx_gather_vg(psi, q, n1, n2, g.inv_d1, g.inv_d2);
            byp = mx_gather_hz(by, q, n1, n2);
            bzp = mx_gather_hy(bz, q, n1, n2);
        }
        const double py = pp.x, pz = pp.y;
        const double a = ga.v, ps = gp.v;
        const double inv1ps = 1.0 / (1 + ps);  // the FP64 1/(1+psi) factor
        const double pys = pst.x, pzs = pst.y;
        // S4 (:373-374) in the first iteration, the half-step momentum after it
        const double p2 = first_iter ? py * py + pz * pz
                                     : 0.25 * (py + pys) * (py + pys) + 0.25 * (pz + pzs) * (pz + pzs);
        const double gm = gamma_of(p2, a, ps, inv1ps);
        const double day = ga.dy, daz = ga.dz, dpy = gp.dy, dpz = gp.dz;
        const double dx = g.dx;
        const double s = dx * inv1ps;
        double pyn = py - s * 0.5 * day;  // P1 (:427-432)
        pyn += s * gm * dpy;
        pyn -= dx * bzp;
        double pzn = pz - s * 0.5 * daz;
        pzn += s * gm * dpz;
        pzn += dx * byp;
        P.ps[p] = make_double2(pyn, pzn);
        // P2 (:439-442)
        double Ym = (q.j + q.fy) + (0.5 * s * pyn) * g.inv_d1;
        double Zm = (q.k + q.fz) + (0.5 * s * pzn) * g.inv_d2;
        qm = pos_from_cells(wrap_cells(Ym, n1), wrap_cells(Zm, n2), n1, n2);
    }
    if (threadIdx.x == 0) {
        sh[4 * (PT / 64)] = qm.j;
        sh[4 * (PT / 64) + 1] = qm.k;
    }
    __syncthreads();
    const int jref = sh[4 * (PT / 64)], kref = sh[4 * (PT / 64) + 1];
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) range_add(qm.j, qm.k, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    const Win win = block_window<CAP>(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_note_overflow(win, ctl);
    win_zero<CAP>(lds, 1, win);
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
    win_flush<CAP>(lds, gl, 1, win, n1, n2);
    // Last, so that these stores do not queue ahead of the particle loads (nothing
    // in this kernel reads the grids).
    {  // jx = jx_bunch(i), jy* = jz* = 0 before the C phase (:473-478, :511-518); rho =
       // rho_bunch(i) before the C phase that deposits rho (the last), chi = 0 before the
       // one that deposits chi (the first)
        const double *rb = si.rho_b3 ? si.rho_b3 + (size_t)slice * si.bstride : nullptr;
        const double *jb = si.jx_b3 ? si.jx_b3 + (size_t)slice * si.bstride : nullptr;
        for (size_t t = (size_t)bx * PT + threadIdx.x; t < nn; t += (size_t)gx * PT) {
            if (init_rho) rho[t] = rb ? rb[t] : 0.0;
            jx[t] = jb ? jb[t] : 0.0;
            if (first_iter) chi[t] = 0.0;
            jyn[t] = 0.0;
            jzn[t] = 0.0;
        }
    }
}
template <int PT, int CAP>
__global__ __launch_bounds__(PT) void k_mx_B(MxPart P, MxGeom g, MxSliceIn si, const double *psi, const double *by,
                                              const double *bz, double *src_acc, double *rho, double *jx,
                                              double *chi, double *jyn, double *jzn, int first_iter, int init_rho,
                                              Ctl *ctl) {
    mx_B<PT, CAP>(P, g, si, psi, by, bz, src_acc, rho, jx, chi, jyn, jzn, first_iter, init_rho, ctl, blockIdx.x,
         blockIdx.y, gridDim.x);
}

// ------------------------------------------------------------------ C phase
// Grids jy*, jz*, jx, and rho (RHO) and chi (CHI).  Within a slice only jx and
// j* feed the next field solve; rho and chi are read only by the slice's
// output.  rho depends on the iteration's p*, so only the last iteration
// deposits it.  chi = -n/(1+psi) at the slice's old positions is the same in
// every iteration (:526), so the first iteration deposits it.  A pass that
// stops in a later iteration then still has it, and no iteration repeats it.
// Second launch bound (waves per SIMD): C fits 80 VGPRs at 6 waves without
// spills (5 otherwise), which hides more of its gather -> deposit -> flush chain
// in batches (256^2 x 10 members: 69.8 -> 65.9 us).  The same bound on B
// (56 VGPRs, 8 waves) made it slower (77.9 -> 83.2 us) and is not used.
template <int PT, int CAP, bool RHO, bool CHI>
__device__ __forceinline__ void mx_C(MxPart P, MxGeom g, MxSliceIn si, const double *psi,
                                              const double *psim, double *rho, double *jx, double *chi,
                                              double *jyn, double *jzn, Ctl *ctl, double *lds, int bx, int wk) {
    constexpr int G = 1 + (RHO ? 1 : 0) + (CHI ? 1 : 0);  // node grids
    ctl += wk;
    MX_SKIP(ctl);
    __shared__ int sh[4 * (PT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2, go = (size_t)wk * nn;
    psi += go;
    psim += go;
    rho += go;
    jx += go;
    chi += go;
    jyn += go;
    jzn += go;
    const double *asq = mx_asq(ctl, si, nn, go);
    const int lp = bx * PT + threadIdx.x;
    const bool valid = lp < mx_np(P, wk);
    const unsigned p = (unsigned)(wk * P.pstride + lp);
    Pos q{0, 0, 0.0, 0.0}, qm{0, 0, 0.0, 0.0};
    double vjy = 0, vjz = 0, vjx = 0, vrho = 0, vchi = 0;
    if (valid) {
        q = load_pos(P, p);
        const double w = P.w[p];
        const double2 pp = P.p[p];
        const double2 pst = P.ps[p];
        const double py = pp.x, pz = pp.y;
        const double pys = pst.x, pzs = pst.y;
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
        const double px = 0.5 * (1 + p2 + a - (1 + ps) * (1 + ps)) * inv1ps;
        // P7 (:523-526) with vx = px/gamma and 1/(1-vx) = gamma/(gamma-px) = gamma/(1+psi)
        // (gamma - px = 1 + psi exactly, from the two closed forms): n vx/(1-vx) =
        // n px/(1+psi), n/(1-vx) = n gamma/(1+psi), -n/((1-vx) gamma) = -n/(1+psi).
        // No divisions, and no cancellation in gamma - px near psi -> -1.
        vjx = w * px * inv1ps;
        if (RHO) vrho = w * gamma_of(p2, a, ps, inv1ps) * inv1ps;
        if (CHI) vchi = -w * inv1ps;
    }
    if (threadIdx.x == 0) {
        sh[4 * (PT / 64)] = q.j;
        sh[4 * (PT / 64) + 1] = q.k;
    }
    __syncthreads();
    const int jref = sh[4 * (PT / 64)], kref = sh[4 * (PT / 64) + 1];
    int ya1, ya2, za1, za2, yb1, yb2, zb1, zb2, yc1, yc2, zc1, zc2;
    double yaf, zaf, ybf, zbf, ycf, zcf;
    ax_half(qm.j, qm.fy, n1, ya1, ya2, yaf);  // jy* on the (dy/2, 0) grid
    ax_node(qm.k, qm.fz, n2, za1, za2, zaf);
    ax_node(qm.j, qm.fy, n1, yb1, yb2, ybf);  // jz* on the (0, dz/2) grid
    ax_half(qm.k, qm.fz, n2, zb1, zb2, zbf);
    ax_node(q.j, q.fy, n1, yc1, yc2, ycf);    // jx, rho, chi on the nodes
    ax_node(q.k, q.fz, n2, zc1, zc2, zcf);
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) {
        range_add(ya1, za1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
        range_add(yb1, zb1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
        range_add(yc1, zc1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    }
    const Win win = block_window<CAP>(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_note_overflow(win, ctl);
    win_zero<CAP>(lds, 2 + G, win);
    __syncthreads();
    double *gc[3];  // node grids in deposit order
    {
        int ig = 0;
        if (RHO) gc[ig++] = rho;
        gc[ig++] = jx;
        if (CHI) gc[ig] = chi;
    }
    {
        double aa[4], ab[4], ac[4 * G];
        cic_w(vjy, yaf, zaf, aa);
        cic_w(vjz, ybf, zbf, ab);
        int ig = 0;
        if (RHO) cic_w(vrho, ycf, zcf, ac + 4 * ig++);
        cic_w(vjx, ycf, zcf, ac + 4 * ig++);
        if (CHI) cic_w(vchi, ycf, zcf, ac + 4 * ig);
        // the quad fold also pays in VALU-bound batches: without it the same-address
        // LDS conflicts cost more (256^2 x 10 members: 61.4 us folded, 86.8 us not)
        bool fa = valid, fb = valid, fc = valid;
        quad_fold<4>(ya1 * n2 + za1, fa, aa);
        quad_fold<4>(yb1 * n2 + zb1, fb, ab);
        quad_fold<4 * G>(yc1 * n2 + zc1, fc, ac);
        if (fa) dep4w(lds, win, jyn, n1, n2, ya1, ya2, za1, za2, aa[0], aa[1], aa[2], aa[3]);
        if (fb) dep4w(lds + CAP, win, jzn, n1, n2, yb1, yb2, zb1, zb2, ab[0], ab[1], ab[2], ab[3]);
        if (fc) dep4w_multi<G, CAP>(lds + 2 * CAP, win, gc, n1, n2, yc1, yc2, zc1, zc2, ac);
    }
    __syncthreads();
    double *gl[2 + G] = {jyn, jzn};
    for (int i = 0; i < G; ++i) gl[2 + i] = gc[i];
    win_flush<CAP>(lds, gl, 2 + G, win, n1, n2);
}
template <int PT, int CAP, bool RHO, bool CHI>
__global__ __launch_bounds__(PT, 6) void k_mx_C(MxPart P, MxGeom g, MxSliceIn si, const double *psi,
                                              const double *psim, double *rho, double *jx, double *chi,
                                              double *jyn, double *jzn, Ctl *ctl) {
    extern __shared__ double ldsd[];
    mx_C<PT, CAP, RHO, CHI>(P, g, si, psi, psim, rho, jx, chi, jyn, jzn, ctl, ldsd, blockIdx.x, blockIdx.y);
}

// -------------------------------------------------- S phase (slice 0, K = 0)
// gamma/px (S4, :368-378) and the rho/jx/chi deposits (S5, :390-397).
template <int PT, int CAP>
__device__ __forceinline__ void mx_S(MxPart P, MxGeom g, MxSliceIn si, const double *psi, double *rho,
                                              double *jx, double *chi, Ctl *ctl, double *lds, int bx, int wk) {
    ctl += wk;
    MX_SKIP(ctl);
    __shared__ int sh[4 * (PT / 64) + 2];
    const int n1 = g.n1, n2 = g.n2;
    const size_t nn = (size_t)n1 * n2, go = (size_t)wk * nn;
    psi += go;
    rho += go;
    jx += go;
    chi += go;
    const double *asq = mx_asq(ctl, si, nn, go);
    const int lp = bx * PT + threadIdx.x;
    const bool valid = lp < mx_np(P, wk);
    const unsigned p = (unsigned)(wk * P.pstride + lp);
    Pos q{0, 0, 0.0, 0.0};
    double vjx = 0, vrho = 0, vchi = 0;
    if (valid) {
        q = load_pos(P, p);
        const double w = P.w[p];
        const double2 pp = P.p[p];
        const double py = pp.x, pz = pp.y;
        const double ps = mx_gather(psi, q, n1, n2);
        const double a = mx_gather(asq, q, n1, n2);
        const double inv1ps = 1.0 / (1 + ps);
        const double gm = gamma_of(py * py + pz * pz, a, ps, inv1ps);
        const double px = 0.5 * (1 + py * py + pz * pz + a - (1 + ps) * (1 + ps)) * inv1ps;
        // S5 (:390-397) with 1/(1-vx) = gamma/(1+psi), as in the C phase
        vjx = w * px * inv1ps;
        vrho = w * gm * inv1ps;
        vchi = -w * inv1ps;
    }
    if (threadIdx.x == 0) {
        sh[4 * (PT / 64)] = q.j;
        sh[4 * (PT / 64) + 1] = q.k;
    }
    __syncthreads();
    const int jref = sh[4 * (PT / 64)], kref = sh[4 * (PT / 64) + 1];
    int y1, y2, z1, z2;
    double yf, zf;
    ax_node(q.j, q.fy, n1, y1, y2, yf);
    ax_node(q.k, q.fz, n2, z1, z2, zf);
    int rmin = INT_MAX, rmax = INT_MIN, cmin = INT_MAX, cmax = INT_MIN;
    if (valid) range_add(y1, z1, jref, kref, n1, n2, rmin, rmax, cmin, cmax);
    const Win win = block_window<CAP>(rmin, rmax, cmin, cmax, jref, kref, n1, n2, sh);
    win_note_overflow(win, ctl);
    win_zero<CAP>(lds, 3, win);
    __syncthreads();
    {
        double ac[12];
        cic_w(vrho, yf, zf, ac);
        cic_w(vjx, yf, zf, ac + 4);
        cic_w(vchi, yf, zf, ac + 8);
        bool act = valid;
        quad_fold<12>(y1 * n2 + z1, act, ac);
        if (act) {
            double *const gc[3] = {rho, jx, chi};
            dep4w_multi<3, CAP>(lds, win, gc, n1, n2, y1, y2, z1, z2, ac);
        }
    }
    __syncthreads();
    double *gl[3] = {rho, jx, chi};
    win_flush<CAP>(lds, gl, 3, win, n1, n2);
}
template <int PT, int CAP>
__global__ __launch_bounds__(PT) void k_mx_S(MxPart P, MxGeom g, MxSliceIn si, const double *psi, double *rho,
                                              double *jx, double *chi, Ctl *ctl) {
    extern __shared__ double ldsd[];
    mx_S<PT, CAP>(P, g, si, psi, rho, jx, chi, ctl, ldsd, blockIdx.x, blockIdx.y);
}

// rho_ion = S(y,z)(-n) (src/System_3d.cpp:318-323), direct atomics (once per pass).
__global__ __launch_bounds__(kPT) void k_mx_ion(MxPart P, MxGeom g, double *rho_ion) {
    const int wk = blockIdx.y;
    const int lp = blockIdx.x * kPT + threadIdx.x;
    if (lp >= mx_np(P, wk)) return;
    const unsigned p = (unsigned)(wk * P.pstride + lp);
    rho_ion += (size_t)wk * g.n1 * g.n2;
    const Pos q = load_pos(P, p);
    int j1, j2, k1, k2;
    double yf, zf;
    ax_node(q.j, q.fy, g.n1, j1, j2, yf);
    ax_node(q.k, q.fz, g.n2, k1, k2, zf);
    scatter4_global(rho_ion, g.n2, j1, j2, k1, k2, yf, zf, -P.w[p]);
}

// ------------------------------------------------------------- output step
// e_x, e_y, e_z of slice i (src/System_3d.cpp:633-661); also hands the slice
// counter to the next step.
// psi of the slice in psi_b[slot], of the slice before in psi_b[1 - slot]; for
// the final output_step(0) the slot follows each member's own stop slice.
__device__ __forceinline__ void mx_out(double *ex, double *ey, double *ez, const double *psi_b0, const double *psi_b1,
                         const double *by, const double *bz, int n1, int n2, double dx, double dy, double dz,
                         int with_ex, Ctl *ctl, int slot, int final_step, int bx, int wk) {
    ctl += wk;
    if (!final_step) {
        MX_SKIP(ctl);
    } else {
        slot = (ctl->stop_acc ? ctl->stop_slice : 0) & 1;
        if (ctl->stop_acc) {
            by = nullptr;
            bz = nullptr;
        }
    }
    const size_t go = (size_t)wk * n1 * n2;
    const double *psi = (slot ? psi_b1 : psi_b0) + go;
    const double *psi_prev = (slot ? psi_b0 : psi_b1) + go;
    ex += go;
    ey += go;
    ez += go;
    if (by) by += go;
    if (bz) bz += go;
    if (!final_step && bx == 0 && threadIdx.x == 0) ctl->slice_of[1 - slot] = ctl->slice_of[slot] - 1;
    const size_t t = (size_t)bx * blockDim.x + threadIdx.x;
    if (t >= (size_t)n1 * n2) return;
    const int j = (int)(t / n2), k = (int)(t % n2);
    if (with_ex) ex[t] = (psi[t] - psi_prev[t]) / dx;
    const size_t tj = j < n1 - 1 ? t + n2 : (size_t)k;
    const size_t tk = k < n2 - 1 ? t + 1 : (size_t)n2 * j;
    ey[t] = -(psi[tj] - psi[t]) / dy + (bz ? bz[t] : 0.0);
    ez[t] = -(psi[tk] - psi[t]) / dz - (by ? by[t] : 0.0);
}
__global__ void k_mx_out(double *ex, double *ey, double *ez, const double *psi_b0, const double *psi_b1,
                         const double *by, const double *bz, int n1, int n2, double dx, double dy, double dz,
                         int with_ex, Ctl *ctl, int slot, int final_step) {
    mx_out(ex, ey, ez, psi_b0, psi_b1, by, bz, n1, n2, dx, dy, dz, with_ex, ctl, slot, final_step, blockIdx.x,
           blockIdx.y);
}

// chi of the slice into the attached envelope (single wake; the slice index
// comes from the control block, so the kernel sits in the captured step).
__global__ void k_mx_chi_store(const double *chi, double *chi3, Ctl *ctl, int slot, size_t nn) {
    MX_SKIP(ctl);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nn) chi3[(size_t)ctl->slice_of[slot] * nn + t] = chi[t];
}

// ------------------------------------------------------- solver functors
// All take (member, which, j, k): member planes sit at + member * nn.
// psi source: jx_bunch(i) - rho_bunch(i) - rho_ion + deposits; clears the accumulator.
struct SrcPsi {
    double *acc;
    const double *rho_ion, *rho_b3, *jx_b3;  // bunch sources: slice i at + i * bstride
    size_t bstride;
    const Ctl *ctl;
    int slot, n2;
    size_t nn;
    __device__ double operator()(int wk, int, int j, int k) const {
        const size_t t = (size_t)j * n2 + k, tw = (size_t)wk * nn + t;
        const int slice = ctl[wk].slice_of[slot];
        const double jb = jx_b3 ? jx_b3[(size_t)slice * bstride + t] : 0.0;
        const double rb = rho_b3 ? rho_b3[(size_t)slice * bstride + t] : 0.0;
        const double v = acc[tw];
        acc[tw] = 0.0;
        return (jb - rb - rho_ion[tw]) + v;
    }
};

// B_y (which = 0, :530-536) and B_z (which = 1, :555-564) sources.
struct SrcB {
    const double *by, *bz, *jx, *jy, *jz, *jyn, *jzn;
    int n1, n2;
    double D, inv_dx, inv_dy, inv_dz;  // the differences are scaled by reciprocals (no divisions)
    __device__ double operator()(int wk, int which, int j, int k) const {
        const size_t go = (size_t)wk * n1 * n2;
        const size_t t = go + (size_t)j * n2 + k;
        if (which == 0) {
            const double jn = k < n2 - 1 ? jx[t + 1] : jx[go + (size_t)j * n2];
            const double djz = (jz[t] - jzn[t]) * inv_dx;
            return -D * by[t] - (jn - jx[t]) * inv_dz + djz;
        }
        const double djy = (jy[t] - jyn[t]) * inv_dx;
        if (j < n1 - 1) return -D * bz[t] + (jx[t + n2] - jx[t]) * inv_dy - djy;
        return -D * bz[go + k] + (jx[go + k] - jx[t]) * inv_dy - djy;
    }
};

// psi finalisation: v/n1/n2, minus the border mean (normalize_potential,
// :30-47) computed from the column pass's partial sums, then the clamp.
struct SinkPsiFinal {
    double *dst;
    const double *bpart;  // per member (stride bstride): [ncb] (P1, P2) pairs then G00
    int ncb, n1, n2;
    double thr;
    double b;
    size_t bstride;
    double scale;  // 1 / (n1 n2): the unnormalised c2r output times this (exact for 2^p sides)
    // Every wavefront sums the partial sums itself (no block barrier); lane 0's
    // result is broadcast, so every lane and every block uses the same value.
    __device__ void begin(int wk) {
        const double *bp = bpart + wk * bstride;
        double s1 = 0.0, s2 = 0.0;
        for (int i = threadIdx.x & 63; i < ncb; i += 64) {
            s1 += bp[2 * i];
            s2 += bp[2 * i + 1];
        }
        for (int off = 32; off > 0; off >>= 1) {
            s1 += __shfl_xor(s1, off);
            s2 += __shfl_xor(s2, off);
        }
        s1 = __shfl(s1, 0);
        s2 = __shfl(s2, 0);
        const double g00 = bp[2 * ncb];
        const double sraw = s1 + (double)n2 * g00 - s2;
        b = sraw / n1 / n2 / (n1 + n2 - 1);
    }
    __device__ void operator()(int wk, int, int j, int k, double v) const {
        double r = v * scale - b;
        if (r < thr) r = thr;
        dst[(size_t)wk * n1 * n2 + (size_t)j * n2 + k] = r;
    }
};

// B store + stop rule, paired: which selects B_y / B_z.
struct SinkBPair {
    double *by, *bz;
    int n1, n2;
    double thr;
    Ctl *ctl;
    int slot;
    double scale;  // 1 / (n1 n2)
    __device__ void begin(int) {}
    __device__ void operator()(int wk, int which, int j, int k, double v) const {
        const double b = v * scale;
        (which == 0 ? by : bz)[(size_t)wk * n1 * n2 + (size_t)j * n2 + k] = b;
        if (fabs(b) > thr || isnan(b)) {
            ctl[wk].stop_acc = 1;
            ctl[wk].stop_slice = ctl[wk].slice_of[slot];
        }
    }
};

// Longitudinal B (extension, no reference counterpart): in the quasistatic
// limit (dx^2 - dt^2 vanishes on functions of t - x) the x component of
// curl curl B = curl j reads lap_perp B_x = dz j_y - dy j_z.  With j_y on
// (dy/2, 0) and j_z on (0, dz/2) the source sits on the cell corners
// (dy/2, dz/2): (j_y[j][k+1] - j_y[j][k])/dz - (j_z[j+1][k] - j_z[j][k])/dy,
// periodic.  Solved with D = 0 (the (0,0) mode dropped: B_x has zero mean).
struct SrcCurl {
    const double *jy, *jz;
    int n1, n2;
    double inv_dy, inv_dz;
    __device__ double operator()(int wk, int, int j, int k) const {
        const size_t go = (size_t)wk * n1 * n2;
        const int kp = k + 1 == n2 ? 0 : k + 1, jp = j + 1 == n1 ? 0 : j + 1;
        const size_t t = go + (size_t)j * n2 + k;
        return (jy[go + (size_t)j * n2 + kp] - jy[t]) * inv_dz - (jz[go + (size_t)jp * n2 + k] - jz[t]) * inv_dy;
    }
};

struct SinkMember {
    double *dst;
    int n1, n2;
    double scale;  // 1 / (n1 n2)
    __device__ void begin(int) {}
    __device__ void operator()(int wk, int, int j, int k, double v) const {
        dst[(size_t)wk * n1 * n2 + (size_t)j * n2 + k] = v * scale;
    }
};

// Column pass of the psi solves, with the partial sums the border mean needs:
// after the inverse column transform G[j][l], the c2r row value at k = 0 is
// sum_l c_l Re G[j][l] (c_0 = c_{n2/2} = 1, else 2) and sum_k x[0][k] = n2 Re G[0][0].
// blockIdx.y is the member (spectra at + member * sstride, partial sums at +
// member * bstride).
__global__ __launch_bounds__(kSolverThreads) void k_cols_psi(double2 *spec, int n1, int nh, int n2, FftAxisDev ax,
                                                             const double *lam1, const double *lam2, int C,
                                                             double *bpart, const Ctl *ctl, size_t sstride,
                                                             size_t bstride) {
    const int wk = blockIdx.y;
    if (solver_skip(ctl + wk, 1)) return;
    spec += wk * sstride;
    bpart += wk * bstride;
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
// Counting sort of the particles by cell, the cells taken 16x16 tile by tile
// and row by row inside a tile: the order of the initial state
// (build_initial_state), so a re-binned block again covers a few rows of one
// tile and the lanes of a quad again share cells.  Particles of one cell come
// out in any order (the deposits are order-independent up to rounding).
constexpr int kTile = 16;
constexpr int kTileBins = kTile * kTile;
__device__ __forceinline__ int sort_bin(uint32_t c, int ntk) {
    const int j = (int)(c >> 16), k = (int)(c & 0xffffu);
    return ((j / kTile) * ntk + k / kTile) * kTileBins + (j % kTile) * kTile + (k % kTile);
}
__global__ void k_sort_hist(MxPart P, unsigned *count, int ntk) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    atomicAdd(&count[sort_bin(P.cell[p], ntk)], 1u);
}

// Exclusive scan of the bin counts: every block scans a chunk of kScanChunk
// bins in place (kScanPer consecutive bins per thread), clears their cursors
// (when given) and writes the chunk's total; the same kernel on one block
// scans the totals; k_scan_add then adds each chunk's offset.
constexpr int kScanTB = 1024, kScanPer = 4, kScanChunk = kScanTB * kScanPer;
__global__ __launch_bounds__(kScanTB) void k_scan_chunks(unsigned *a, unsigned *cursor, unsigned *totals, int n) {
    __shared__ unsigned wsum[kScanTB / 64];
    const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
    unsigned v[kScanPer], s = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        v[i] = base + i < n ? a[base + i] : 0u;
        s += v[i];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned x = s;  // inclusive scan over the wavefront
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    unsigned run = x - s;
    for (int i = 0; i < wv; ++i) run += wsum[i];
    if (threadIdx.x == kScanTB - 1) totals[blockIdx.x] = run + s;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        if (base + i < n) {
            a[base + i] = run;
            if (cursor) cursor[base + i] = 0;
        }
        run += v[i];
    }
}
__global__ void k_scan_add(unsigned *count, const unsigned *offs, int n) {
    const int b = blockIdx.x * kPT + threadIdx.x;
    if (b < n) count[b] += offs[b / kScanChunk];
}

__global__ void k_sort_scatter(MxPart P, MxPart Q, const unsigned *offset, unsigned *cursor, const int *perm,
                               int *perm2, int ntk) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    const uint32_t c = P.cell[p];
    const int t = sort_bin(c, ntk);
    const unsigned d = offset[t] + atomicAdd(&cursor[t], 1u);
    if (d >= (unsigned)P.np) return;
    Q.cell[d] = c;
    Q.off[d] = P.off[p];
    Q.p[d] = P.p[p];
    Q.w[d] = P.w[p];
    Q.ps[d] = P.ps[p];
    perm2[d] = perm[p];
}

__global__ void k_sort_copyback(MxPart P, MxPart Q, int *perm, const int *perm2) {
    const int p = blockIdx.x * kPT + threadIdx.x;
    if (p >= P.np) return;
    P.cell[p] = Q.cell[p];
    P.off[p] = Q.off[p];
    P.p[p] = Q.p[p];
    P.w[p] = Q.w[p];
    P.ps[p] = Q.ps[p];
    perm[p] = perm2[p];
}

// ----------------------------------------------------------- group launch
// A group advances several batches (each with its own transverse grid) in
// lockstep, and every phase of the slice step is one launch over all of them:
// the flat block index walks the batches' own launch grids one after another
// (batch b owns blocks [start[b], start[b+1]), laid out as gx[b] blocks per
// member row), and each block reads its batch's arguments from a device table.
// The solver phases run the register-resident FFT of the batch's grid side
// (a uniform switch) with one launch shape for all sides, 64 threads per
// block, so the LDS buffers are one dynamic allocation sized for the largest
// transform group (sides 3 * 2^p: eight 96-point or four 192-point transforms).
constexpr int kGroupMax = 16;
constexpr int kGrpTB = 64;
// LDS of the two transform families per 64-thread block: T * PADN = 64 E / N *
// 17 N / 16 = 68 E complex values per exchange buffer (E = 4 for sides 2^p,
// kWfE3 for 3 * 2^p).  The families launch separately: one kernel holding both
// would give every block the registers and LDS of the 3 * 2^p transforms.
constexpr int grp_lds(int fam) { return 68 * (fam ? kWfE3 : 4); }

struct GrpMap {
    int nb;                      // entries
    int start[kGroupMax + 1];    // first block of entry i
    int gx[kGroupMax];           // x extent of entry i's own launch grid
    int bi[kGroupMax];           // batch (table index) of entry i
};
__device__ __forceinline__ void grp_locate(const GrpMap &m, int &b, int &bx, int &by) {
    const int f = (int)blockIdx.x;
    int i = 0;
    while (i + 1 < m.nb && f >= m.start[i + 1]) ++i;
    b = m.bi[i];
    const int r = f - m.start[i];
    by = r / m.gx[i];
    bx = r - by * m.gx[i];
}

struct GArgsA {
    MxPart P;
    MxGeom g;
    MxSliceIn si;
    const double *psim;
    double *src_acc, *asq2;
    const LaserGauss *Ls;
    int laser;
    double *rho, *jx, *chi;
    Ctl *ctl;
};
struct GArgsB {
    MxPart P;
    MxGeom g;
    MxSliceIn si;
    const double *psi, *by, *bz;
    double *src_acc, *rho, *jx, *chi, *jyn, *jzn;
    Ctl *ctl;
};
struct GArgsC {
    MxPart P;
    MxGeom g;
    MxSliceIn si;
    const double *psi, *psim;
    double *rho, *jx, *chi, *jyn, *jzn;
    Ctl *ctl;
};
struct GArgsOut {
    double *ex, *ey, *ez;
    const double *psi_b0, *psi_b1, *by, *bz;
    int n1, n2;
    double dx, dy, dz;
    Ctl *ctl;
};
template <class Src>
struct GArgsRowsF {
    Src src;
    double2 *spec0, *spec1;
    int n1, n2;
    const double2 *tw2;
    const Ctl *ctl;
    size_t sstride;
};
struct GArgsCols {
    double2 *spec0, *spec1;
    int n1, nh, n2;
    const double2 *tw1;
    const double *lam1, *lam2;
    double D;
    double *bpart;
    const Ctl *ctl;
    size_t sstride, bstride;
};
template <class Sink>
struct GArgsRowsI {
    Sink sink;
    const double2 *spec0, *spec1;
    int n1, n2;
    const double2 *tw2;
    const Ctl *ctl;
    size_t sstride;
};

// The group kernels read their arguments from the table (not the kernel
// argument segment): left alone the compiler holds more of them in vector
// registers than the single-batch kernels do (B 88 VGPRs and 5 waves per SIMD
// against 70 and 7, A 68 against 50).  The bounds restore 7 and 8 waves at
// 256 threads without scratch (B 66, A 64 VGPRs).
template <int PT, int CAP>
__global__ __launch_bounds__(PT, 8) void kg_mx_A(GrpMap m, const GArgsA *__restrict__ t, int init_rho, int advance) {
    int b, bx, wk;
    grp_locate(m, b, bx, wk);
    const GArgsA &a = t[b];
    mx_A<PT, CAP>(a.P, a.g, a.si, a.psim, a.src_acc, a.asq2, a.Ls, a.laser, a.rho, a.jx, a.chi, init_rho, advance, a.ctl, bx, wk,
         m.gx[b]);
}
template <int PT, int CAP>
__global__ __launch_bounds__(PT, 7) void kg_mx_B(GrpMap m, const GArgsB *__restrict__ t, int first_iter, int init_rho) {
    int b, bx, wk;
    grp_locate(m, b, bx, wk);
    const GArgsB &a = t[b];
    mx_B<PT, CAP>(a.P, a.g, a.si, a.psi, a.by, a.bz, a.src_acc, a.rho, a.jx, a.chi, a.jyn, a.jzn, first_iter, init_rho, a.ctl,
         bx, wk, m.gx[b]);
}
template <int PT, int CAP, bool RHO, bool CHI>
__global__ __launch_bounds__(PT, 6) void kg_mx_C(GrpMap m, const GArgsC *__restrict__ t) {
    extern __shared__ double ldsd[];
    int b, bx, wk;
    grp_locate(m, b, bx, wk);
    const GArgsC &a = t[b];
    mx_C<PT, CAP, RHO, CHI>(a.P, a.g, a.si, a.psi, a.psim, a.rho, a.jx, a.chi, a.jyn, a.jzn, a.ctl, ldsd, bx, wk);
}
// S phase: the C table carries the same pointers (psi, rho, jx, chi).
template <int PT, int CAP>
__global__ __launch_bounds__(PT) void kg_mx_S(GrpMap m, const GArgsC *__restrict__ t) {
    extern __shared__ double ldsd[];
    int b, bx, wk;
    grp_locate(m, b, bx, wk);
    const GArgsC &a = t[b];
    mx_S<PT, CAP>(a.P, a.g, a.si, a.psi, a.rho, a.jx, a.chi, a.ctl, ldsd, bx, wk);
}
__global__ void kg_mx_out(GrpMap m, const GArgsOut *__restrict__ t, int with_ex, int slot, int final_step) {
    int b, bx, wk;
    grp_locate(m, b, bx, wk);
    const GArgsOut &a = t[b];
    mx_out(a.ex, a.ey, a.ez, a.psi_b0, a.psi_b1, a.by, a.bz, a.n1, a.n2, a.dx, a.dy, a.dz, with_ex, a.ctl, slot,
           final_step, bx, wk);
}

// Grid sides a group launch takes: the register-resident lengths whose
// transforms fit a 64-thread block with register twiddles.
inline bool grp_side_ok(int n) {
    switch (n) {
        case 16: case 32: case 48: case 64: case 96: case 128: case 192: case 256: return true;
        default: return false;
    }
}
#define QP_GRP_SIDES_P2(F) F(16) F(32) F(64) F(128) F(256)
#define QP_GRP_SIDES_P3(F) F(48) F(96) F(192)
// Elements per lane of a side's transform (WfCfg<N>::E at run time).
inline int grp_elems(int n) { return n % 3 == 0 ? kWfE3 : 4; }

template <class Src, int FAM>
__global__ __launch_bounds__(kGrpTB) void kg_rows_fwd(GrpMap m, const GArgsRowsF<Src> *__restrict__ t, int skip_mode,
                                                      int nspec) {
    extern __shared__ double2 gl2[];
    int b, bx, by;
    grp_locate(m, b, bx, by);
    const GArgsRowsF<Src> &a = t[b];
    double2 *sb0 = gl2, *sb1 = gl2 + grp_lds(FAM);
#define QP_CASE(N)                                                                                                 \
    case N:                                                                                                        \
        wrows_fwd_body<N, WfCfg<N>::E, kGrpTB, Src>(a.src, a.spec0, a.spec1, a.n1, a.tw2, a.ctl, skip_mode, nspec, \
                                                    a.sstride, bx, by, sb0, sb1, nullptr);                         \
        break;
    if constexpr (FAM == 0) {
        switch (a.n2) { QP_GRP_SIDES_P2(QP_CASE) default: break; }
    } else {
        switch (a.n2) { QP_GRP_SIDES_P3(QP_CASE) default: break; }
    }
#undef QP_CASE
}
// The border partial sums are indexed by the block's position in its batch's
// own column grid (bx) and that grid's size.
template <int FAM>
__global__ __launch_bounds__(kGrpTB) void kg_cols(GrpMap m, const GArgsCols *__restrict__ t, int skip_mode, int nspec) {
    extern __shared__ double2 gl2[];
    int b, bx, by;
    grp_locate(m, b, bx, by);
    const GArgsCols &a = t[b];
    const int tpb = kGrpTB * (FAM ? kWfE3 : 4) / a.n1;   // transforms per block
    const int gx = (a.nh + tpb - 1) / tpb;
    double2 *sb0 = gl2, *sb1 = gl2 + grp_lds(FAM);
#define QP_CASE(N)                                                                                                  \
    case N:                                                                                                         \
        wcols_body<N, WfCfg<N>::E, kGrpTB>(a.spec0, a.spec1, a.nh, a.n2, a.tw1, a.lam1, a.lam2, a.D, a.bpart, a.ctl, \
                                           skip_mode, nspec, a.sstride, a.bstride, bx, by, gx, sb0, sb1, nullptr);  \
        break;
    if constexpr (FAM == 0) {
        switch (a.n1) { QP_GRP_SIDES_P2(QP_CASE) default: break; }
    } else {
        switch (a.n1) { QP_GRP_SIDES_P3(QP_CASE) default: break; }
    }
#undef QP_CASE
}
template <class Sink, int FAM>
__global__ __launch_bounds__(kGrpTB) void kg_rows_inv(GrpMap m, const GArgsRowsI<Sink> *__restrict__ t, int skip_mode,
                                                      int nspec) {
    extern __shared__ double2 gl2[];
    int b, bx, by;
    grp_locate(m, b, bx, by);
    const GArgsRowsI<Sink> &a = t[b];
    double2 *sb0 = gl2, *sb1 = gl2 + grp_lds(FAM);
#define QP_CASE(N)                                                                                                 \
    case N:                                                                                                        \
        wrows_inv_body<N, WfCfg<N>::E, kGrpTB, Sink, true>(a.spec0, a.spec1, a.sink, a.n1, a.tw2, a.ctl, skip_mode, \
                                                           nspec, a.sstride, bx, by, sb0, sb1, nullptr);           \
        break;
    if constexpr (FAM == 0) {
        switch (a.n2) { QP_GRP_SIDES_P2(QP_CASE) default: break; }
    } else {
        switch (a.n2) { QP_GRP_SIDES_P3(QP_CASE) default: break; }
    }
#undef QP_CASE
}

// ------------------------------------------------------------------ host
// A batch of nw members (nw = 1 for a single wake).  Member m's particles take
// slots [m * pstride, m * pstride + np_m) of every particle array; its grids,
// spectra, partial sums and control block sit at member offsets as well.
struct MixedState {
    MxPart P{}, Q{}, I{};   // live, re-binning scratch, initial state of a pass
    int *perm = nullptr, *perm2 = nullptr, *perm0 = nullptr;  // local particle index per slot (-1: padding)
    int *npw = nullptr;     // particles per member (device)
    int pstride = 0;
    bool init_valid = false;
    unsigned *count = nullptr, *cursor = nullptr, *totals = nullptr;  // re-binning: per bin, per scan chunk
    double *bpart = nullptr;
    size_t bstride = 0;     // partial sums per member
    double *jbuf[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [parity][y/z], nw planes each
    std::vector<int> h_perm;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    unsigned graph_epoch = 0;  // Wake::inputs_epoch the graphs were captured at
    int graph_pt = 0, graph_cap = 0;  // particle tile the graphs were captured with
    // streamed output (async_output): per-slice snapshots of the output planes
    // and control blocks in two device slots (slice parity), copied on a copy
    // stream into a ring of pinned host slots that a writer thread hands to the
    // callback
    double *snap_dev = nullptr;        // 2 slots
    double *snap_host = nullptr;       // kHostSlots slots (pinned)
    size_t snap_words = 0;             // doubles per slot
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_done[2] = {nullptr, nullptr}, ev_copied[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_host;
    int ncb = 0;
    size_t cap = 0;
};

static MixedState *mstate(Wake *w) { return static_cast<MixedState *>(w->mixed); }

// Host <-> device copy ordered on the wake's stream, waited for.  Plain
// hipMemcpy would go through the legacy stream, which HIP refuses while another
// host thread captures a graph on its own stream (batches run from several
// threads, each capturing its slice steps).
static hipError_t copy_wait(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

static void free_part(MxPart &P) {
    void *arrs[] = {P.cell, P.off, P.p, P.w, P.ps};
    for (void *a : arrs)
        if (a) (void)hipFree(a);
    P = MxPart{};
}

static int alloc_part(MxPart &P, size_t n) {
    n = n ? n : 1;
    QP_HIP_CHECK(hipMalloc(&P.cell, 4 * n));
    QP_HIP_CHECK(hipMalloc(&P.off, 8 * n));
    QP_HIP_CHECK(hipMalloc(&P.p, 16 * n));
    QP_HIP_CHECK(hipMalloc(&P.w, 8 * n));
    QP_HIP_CHECK(hipMalloc(&P.ps, 16 * n));
    return QP_OK;
}

void wake_destroy_mixed(Wake *w) {
    MixedState *m = mstate(w);
    if (!m) return;
    free_part(m->P);
    free_part(m->Q);
    free_part(m->I);
    void *arrs[] = {m->perm,       m->perm2,      m->perm0,      m->npw,        m->count,     m->cursor,
                    m->totals,     m->bpart,      m->jbuf[0][0], m->jbuf[0][1], m->jbuf[1][0], m->jbuf[1][1]};
    for (void *a : arrs)
        if (a) (void)hipFree(a);
    for (auto &g : m->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (m->snap_dev) (void)hipFree(m->snap_dev);
    if (m->snap_host) (void)hipHostFree(m->snap_host);
    for (hipEvent_t e : {m->ev_done[0], m->ev_done[1], m->ev_copied[0], m->ev_copied[1]})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : m->ev_host)
        if (e) (void)hipEventDestroy(e);
    if (m->copy_stream) (void)hipStreamDestroy(m->copy_stream);
    delete m;
    w->mixed = nullptr;
}

// Slots per member: the largest member rounded up to whole blocks.
static int member_stride(const Wake *w) {
    size_t mx = 0;
    for (const WakeMember &mb : w->mem) mx = std::max(mx, mb.np);
    const size_t st = ((mx + kPT - 1) / kPT) * kPT;
    return (int)std::max<size_t>(st, kPT);
}

static int ensure_state(Wake *w) {
    MixedState *m = mstate(w);
    const size_t nn = w->nn(), nw = (size_t)w->nw;
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
        m->bstride = 2 * (size_t)ncb + 2;
        QP_HIP_CHECK(hipMalloc(&m->bpart, sizeof(double) * m->bstride * nw));
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) QP_HIP_CHECK(hipMalloc(&m->jbuf[a][b], sizeof(double) * nn * nw));
        const int ntk = (w->p.nz + kTile - 1) / kTile, ntj = (w->p.ny + kTile - 1) / kTile;
        const size_t nbins = (size_t)ntk * ntj * kTileBins;
        QP_HIP_CHECK(hipMalloc(&m->count, sizeof(unsigned) * nbins));
        QP_HIP_CHECK(hipMalloc(&m->cursor, sizeof(unsigned) * nbins));
        QP_HIP_CHECK(hipMalloc(&m->totals, sizeof(unsigned) * ((nbins + kScanChunk - 1) / kScanChunk + 1)));
        QP_HIP_CHECK(hipMalloc(&m->npw, sizeof(int) * nw));
    }
    const int pstride = member_stride(w);
    const size_t need = nw * (size_t)pstride;
    if (m->cap < need || !m->P.cell || m->pstride != pstride) {
        free_part(m->P);
        free_part(m->Q);
        free_part(m->I);
        if (m->perm) (void)hipFree(m->perm);
        if (m->perm2) (void)hipFree(m->perm2);
        if (m->perm0) (void)hipFree(m->perm0);
        QP_TRY(alloc_part(m->P, need));
        QP_TRY(alloc_part(m->Q, need));
        QP_TRY(alloc_part(m->I, need));
        QP_HIP_CHECK(hipMalloc(&m->perm, sizeof(int) * need));
        QP_HIP_CHECK(hipMalloc(&m->perm2, sizeof(int) * need));
        QP_HIP_CHECK(hipMalloc(&m->perm0, sizeof(int) * need));
        m->cap = need;
        m->init_valid = false;
        for (auto &g : m->gexec)
            if (g) (void)hipGraphExecDestroy(g);
        m->gexec[0] = m->gexec[1] = nullptr;
    }
    m->pstride = pstride;
    std::vector<int> hn(nw);
    for (size_t i = 0; i < nw; ++i) hn[i] = (int)w->mem[i].np;
    QP_HIP_CHECK(copy_wait(m->npw, hn.data(), sizeof(int) * nw, hipMemcpyHostToDevice, w->stream()));
    for (MxPart *q : {&m->P, &m->Q, &m->I}) {
        q->np = (int)w->mem[0].np;
        q->pstride = pstride;
        q->npw = m->npw;
    }
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
// 16x16-cell tile, then by cell) into the pass's initial-state buffers; every
// pass then starts with device-to-device copies, as the reference
// re-initialises its particles each pass (src/System_3d.cpp:262).
static int build_initial_state(Wake *w, MixedState *m) {
    const int n1 = w->p.ny, n2 = w->p.nz;
    const int ntk = (n2 + kTile - 1) / kTile;
    const size_t total = (size_t)w->nw * m->pstride;
    std::vector<uint32_t> c2(total, 0), oyz(2 * total, 0);
    std::vector<double> pyz(2 * total, 0.0);
    std::vector<double> wt(total, 0.0);
    m->h_perm.assign(total, -1);
    for (int mi = 0; mi < w->nw; ++mi) {
        const WakeMember &mb = w->mem[mi];
        const size_t np = mb.np, base = (size_t)mi * m->pstride;
        std::vector<uint32_t> cell(np), oy(np), oz(np);
        std::vector<long long> key(np);
        for (size_t i = 0; i < np; ++i) {
            int j, k;
            to_cells(mb.h_y[i], w->p.dy, n1, j, oy[i]);
            to_cells(mb.h_z[i], w->p.dz, n2, k, oz[i]);
            cell[i] = ((uint32_t)j << 16) | (uint32_t)k;
            const long long tile = (long long)(j / kTile) * ntk + (k / kTile);
            key[i] = (tile * n1 + j) * (long long)n2 + k;
        }
        std::vector<int> ord(np);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
        for (size_t s = 0; s < np; ++s) {
            const int i = ord[s];
            c2[base + s] = cell[i];
            oyz[2 * (base + s)] = oy[i];
            oyz[2 * (base + s) + 1] = oz[i];
            pyz[2 * (base + s)] = mb.h_py[i];
            pyz[2 * (base + s) + 1] = mb.h_pz[i];
            wt[base + s] = mb.h_w[i];
            m->h_perm[base + s] = i;
        }
    }
    hipStream_t s = w->stream();
    QP_HIP_CHECK(copy_wait(m->I.cell, c2.data(), 4 * total, hipMemcpyHostToDevice, s));
    QP_HIP_CHECK(copy_wait(m->I.off, oyz.data(), 8 * total, hipMemcpyHostToDevice, s));
    QP_HIP_CHECK(copy_wait(m->I.p, pyz.data(), 16 * total, hipMemcpyHostToDevice, s));
    QP_HIP_CHECK(copy_wait(m->I.w, wt.data(), 8 * total, hipMemcpyHostToDevice, s));
    QP_HIP_CHECK(copy_wait(m->perm0, m->h_perm.data(), sizeof(int) * total, hipMemcpyHostToDevice, s));
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
    const size_t n = (size_t)w->nw * m->pstride;
    hipStream_t s = w->stream();
    QP_HIP_CHECK(hipMemcpyAsync(m->P.cell, m->I.cell, 4 * n, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.off, m->I.off, 8 * n, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.p, m->I.p, 16 * n, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->P.w, m->I.w, 8 * n, hipMemcpyDeviceToDevice, s));
    // py_next = pz_next = 0 (src/System_3d.cpp:740-741): the first slice reads no
    // increment, and with K = 0 the commit writes p = 0 directly
    QP_HIP_CHECK(hipMemsetAsync(m->P.ps, 0, 16 * n, s));
    QP_HIP_CHECK(hipMemcpyAsync(m->perm, m->perm0, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    return QP_OK;
}

int wake_download_mixed(Wake *w, int member, double *y, double *z, double *py, double *pz) {
    MixedState *m = mstate(w);
    if (!m) {
        set_error("mixed state not initialised");
        return QP_ERR_STATE;
    }
    const size_t np = w->mem[member].np, n = (size_t)m->pstride, base = (size_t)member * m->pstride;
    QP_HIP_CHECK(hipStreamSynchronize(w->stream()));
    std::vector<uint32_t> c(n), oyz(2 * n);
    std::vector<double> pyz(2 * n);
    std::vector<int> perm(n);
    hipStream_t s = w->stream();
    QP_HIP_CHECK(copy_wait(c.data(), m->P.cell + base, 4 * n, hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(copy_wait(oyz.data(), m->P.off + base, 8 * n, hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(copy_wait(pyz.data(), m->P.p + base, 16 * n, hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(copy_wait(perm.data(), m->perm + base, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    for (size_t s = 0; s < n; ++s) {
        const int i = perm[s];
        if (i < 0 || (size_t)i >= np) continue;
        if (y) y[i] = ((double)(c[s] >> 16) + oyz[2 * s] * kInvTwo32) * w->p.dy;
        if (z) z[i] = ((double)(c[s] & 0xffffu) + oyz[2 * s + 1] * kInvTwo32) * w->p.dz;
        if (py) py[i] = pyz[2 * s];
        if (pz) pz[i] = pyz[2 * s + 1];
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
    int nw, pb, gb;
    size_t sstride;
    size_t ldsS;
    int pt, cap;   // particle tile: threads per block, LDS window nodes per grid
};

// The instantiated particle tiles (threads per block, window nodes per grid).
// The LDS of the C phase is (2 + grids) x CAP doubles: 25.6 KB at (256, 640),
// 40 KB at (256, 1024).
constexpr int kTiles[][2] = {{256, 640}, {128, 640}, {256, 1024}, {128, 384}};
constexpr int kNumTiles = 4;

inline bool tile_valid(int pt, int cap) {
    for (const auto &t : kTiles)
        if (t[0] == pt && t[1] == cap) return true;
    return false;
}

// f(integral_constant<PT>, integral_constant<CAP>) for the tile (pt, cap).
template <class F>
int with_tile(int pt, int cap, F &&f) {
    if (pt == 256 && cap == 640) return f(std::integral_constant<int, 256>{}, std::integral_constant<int, 640>{});
    if (pt == 128 && cap == 640) return f(std::integral_constant<int, 128>{}, std::integral_constant<int, 640>{});
    if (pt == 256 && cap == 1024) return f(std::integral_constant<int, 256>{}, std::integral_constant<int, 1024>{});
    if (pt == 128 && cap == 384) return f(std::integral_constant<int, 128>{}, std::integral_constant<int, 384>{});
    set_error("particle tile (" + std::to_string(pt) + ", " + std::to_string(cap) + ") not instantiated");
    return QP_ERR_ARG;
}

#define MLAUNCH(kern, grid, block, lds, ...)                                            \
    do {                                                                                \
        hipLaunchKernelGGL(kern, grid, dim3(block), lds, s, __VA_ARGS__);               \
        QP_HIP_CHECK(hipGetLastError());                                                \
    } while (0)

int psi_solve(MxRun &R, hipStream_t s, int slot, double *dst) {
    Wake *w = R.w;
    const Solver2d &S = w->S;
    MixedState *m = R.m;
    const int nw = R.nw;
    SrcPsi src{w->src, w->rho_ion, w->rho_b_src(), w->jx_b_src(), w->bsrc_stride(), w->ctl, slot, S.n2, R.nn};
    SinkPsiFinal snk{dst, m->bpart, m->ncb, S.n1, S.n2, w->p.psi_threshold - 1, 0.0, m->bstride,
                     1.0 / ((double)S.n1 * S.n2)};
    if (S.wave) {
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_fwd(S, src, S.spec[0], S.spec[0], 1, s, (const Ctl *)w->ctl, 1, nw));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_cols(S, 0.0, 1, m->bpart, s, (const Ctl *)w->ctl, 1, nw, m->bstride));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_inv(S, (const double2 *)S.spec[0], (const double2 *)S.spec[0], snk, 1, s,
                                (const Ctl *)w->ctl, 1, nw));
        prof_end(w, "solve");
        return QP_OK;
    }
    prof_begin(w, "solve");
    MLAUNCH(k_rows_fwd<SrcPsi>, dim3(rows_blocks(S), nw), kSolverThreads, rows_lds(S), src, S.spec[0], S.spec[0],
            S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)w->ctl, 1, 1, R.sstride);
    prof_end(w, "solve");
    prof_begin(w, "solve");
    MLAUNCH(k_cols_psi, dim3(m->ncb, nw), kSolverThreads, cols_lds(S), S.spec[0], S.n1, S.nh, S.n2, S.ax1.dev,
            (const double *)S.d_lam1, (const double *)S.d_lam2, S.cols_per_block, m->bpart, (const Ctl *)w->ctl,
            R.sstride, m->bstride);
    prof_end(w, "solve");
    prof_begin(w, "solve");
    MLAUNCH(k_rows_inv<SinkPsiFinal>, dim3(rows_blocks(S), nw), kSolverThreads, rows_lds(S),
            (const double2 *)S.spec[0], (const double2 *)S.spec[0], snk, S.n1, S.n2, S.nh, S.ax2.dev,
            S.rows_per_block, (const Ctl *)w->ctl, 1, 1, R.sstride);
    prof_end(w, "solve");
    return QP_OK;
}

int b_solve(MxRun &R, hipStream_t s, int slot) {
    Wake *w = R.w;
    const Solver2d &S = w->S;
    MixedState *m = R.m;
    const qp_wake_params &p = w->p;
    const int nw = R.nw;
    double *jy = m->jbuf[1 - slot][0], *jz = m->jbuf[1 - slot][1];
    double *jyn = m->jbuf[slot][0], *jzn = m->jbuf[slot][1];
    SrcB src{w->by, w->bz, w->jx, jy, jz, jyn, jzn, S.n1, S.n2, p.D, 1.0 / p.dx, 1.0 / p.dy, 1.0 / p.dz};
    SinkBPair snk{w->by, w->bz, S.n1, S.n2, p.b_threshold, w->ctl, slot, 1.0 / ((double)S.n1 * S.n2)};
    if (S.wave) {
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_fwd(S, src, S.spec[0], S.spec[1], 2, s, (const Ctl *)w->ctl, 2, nw));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_cols(S, p.D, 2, nullptr, s, (const Ctl *)w->ctl, 2, nw));
        prof_end(w, "solve");
        prof_begin(w, "solve");
        QP_TRY(wlaunch_rows_inv(S, (const double2 *)S.spec[0], (const double2 *)S.spec[1], snk, 2, s,
                                (const Ctl *)w->ctl, 2, nw));
        prof_end(w, "solve");
        return QP_OK;
    }
    prof_begin(w, "solve");
    MLAUNCH(k_rows_fwd<SrcB>, dim3(rows_blocks(S), 2 * nw), kSolverThreads, rows_lds(S), src, S.spec[0], S.spec[1],
            S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)w->ctl, 2, 2, R.sstride);
    prof_end(w, "solve");
    prof_begin(w, "solve");
    QP_TRY(solve_cols_batch(S, p.D, 2, nw, s, w->ctl, 2));
    prof_end(w, "solve");
    prof_begin(w, "solve");
    MLAUNCH(k_rows_inv<SinkBPair>, dim3(rows_blocks(S), 2 * nw), kSolverThreads, rows_lds(S),
            (const double2 *)S.spec[0], (const double2 *)S.spec[1], snk, S.n1, S.n2, S.nh, S.ax2.dev,
            S.rows_per_block, (const Ctl *)w->ctl, 2, 2, R.sstride);
    prof_end(w, "solve");
    return QP_OK;
}

// B_x of the slice from the output currents j_y, j_z (buffer of parity 1 - slot).
int bx_solve(MxRun &R, hipStream_t s, int slot) {
    Wake *w = R.w;
    const Solver2d &S = w->S;
    MixedState *m = R.m;
    const qp_wake_params &p = w->p;
    const int nw = R.nw;
    SrcCurl src{m->jbuf[1 - slot][0], m->jbuf[1 - slot][1], S.n1, S.n2, 1.0 / p.dy, 1.0 / p.dz};
    SinkMember snk{w->bx, S.n1, S.n2, 1.0 / ((double)S.n1 * S.n2)};
    if (S.wave) {
        QP_TRY(wlaunch_rows_fwd(S, src, S.spec[0], S.spec[0], 1, s, (const Ctl *)w->ctl, 1, nw));
        QP_TRY(wlaunch_cols(S, 0.0, 1, nullptr, s, (const Ctl *)w->ctl, 1, nw));
        QP_TRY(wlaunch_rows_inv(S, (const double2 *)S.spec[0], (const double2 *)S.spec[0], snk, 1, s,
                                (const Ctl *)w->ctl, 1, nw));
        return QP_OK;
    }
    MLAUNCH(k_rows_fwd<SrcCurl>, dim3(rows_blocks(S), nw), kSolverThreads, rows_lds(S), src, S.spec[0], S.spec[0],
            S.n1, S.n2, S.nh, S.ax2.dev, S.rows_per_block, (const Ctl *)w->ctl, 1, 1, R.sstride);
    QP_TRY(solve_cols_batch(S, 0.0, 1, nw, s, w->ctl, 1));
    MLAUNCH(k_rows_inv<SinkMember>, dim3(rows_blocks(S), nw), kSolverThreads, rows_lds(S),
            (const double2 *)S.spec[0], (const double2 *)S.spec[0], snk, S.n1, S.n2, S.nh, S.ax2.dev,
            S.rows_per_block, (const Ctl *)w->ctl, 1, 1, R.sstride);
    return QP_OK;
}

// Enqueue the work of slice i (everything between two output steps) for every member.
template <int PT, int CAP>
int enqueue_slice_t(MxRun &R, hipStream_t s, int i) {
    Wake *w = R.w;
    MixedState *m = R.m;
    const qp_wake_params &p = w->p;
    const int slot = i & 1;
    const bool first = (i == p.nx - 1);
    const bool s_phase = (i == 0) || (p.K == 0);
    const MxSliceIn &si = R.si[slot];
    double *psi = w->psi_b[slot];
    const int laser = (w->any_laser() && !w->asq3()) ? 1 : 0;
    const dim3 pg(R.pb, R.nw);
    if (w->bunch) QP_TRY(bunch_slice_sources(w->bunch, (const Ctl *)w->ctl, slot, 0, s));
    prof_begin(w, "push_A");
    MLAUNCH((k_mx_A<PT, CAP>), pg, PT, 0, m->P, R.g, si, (const double *)w->psim, w->src, w->asq2,
            (const LaserGauss *)w->d_laser, laser, w->rho, w->jx, w->chi, s_phase ? 1 : 0,
            first ? 0 : (p.K == 0 ? 2 : 1), w->ctl);
    prof_end(w, "push_A");
    QP_TRY(psi_solve(R, s, slot, psi));
    if (s_phase) {
        prof_begin(w, "push_S");
        MLAUNCH((k_mx_S<PT, CAP>), pg, PT, R.ldsS, m->P, R.g, si, (const double *)psi, w->rho, w->jx, w->chi, w->ctl);
        prof_end(w, "push_S");
        if (p.longitudinal_b) QP_TRY(bx_solve(R, s, slot));
        if (w->env)
            MLAUNCH(k_mx_chi_store, dim3(R.gb), 256, 0, (const double *)w->chi, w->env->chi3, w->ctl, slot, R.nn);
        if (i == 0) return QP_OK;
    } else {
        for (int it = 0; it < p.K; ++it) {
            prof_begin(w, "push_B");
            const bool first_it = it == 0, last_it = it == p.K - 1;
            MLAUNCH((k_mx_B<PT, CAP>), pg, PT, 0, m->P, R.g, si, (const double *)psi, (const double *)w->by,
                    (const double *)w->bz, w->src, w->rho, w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1],
                    first_it ? 1 : 0, last_it ? 1 : 0, w->ctl);
            prof_end(w, "push_B");
            QP_TRY(psi_solve(R, s, slot, w->psim));
            prof_begin(w, "push_C");
            const size_t ldsc = sizeof(double) * CAP * (3 + (first_it ? 1 : 0) + (last_it ? 1 : 0));
            if (first_it && last_it)
                MLAUNCH((k_mx_C<PT, CAP, true, true>), pg, PT, ldsc, m->P, R.g, si, (const double *)psi,
                        (const double *)w->psim, w->rho, w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1], w->ctl);
            else if (first_it)
                MLAUNCH((k_mx_C<PT, CAP, false, true>), pg, PT, ldsc, m->P, R.g, si, (const double *)psi,
                        (const double *)w->psim, w->rho, w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1], w->ctl);
            else if (last_it)
                MLAUNCH((k_mx_C<PT, CAP, true, false>), pg, PT, ldsc, m->P, R.g, si, (const double *)psi,
                        (const double *)w->psim, w->rho, w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1], w->ctl);
            else
                MLAUNCH((k_mx_C<PT, CAP, false, false>), pg, PT, ldsc, m->P, R.g, si, (const double *)psi,
                        (const double *)w->psim, w->rho, w->jx, w->chi, m->jbuf[slot][0], m->jbuf[slot][1], w->ctl);
            prof_end(w, "push_C");
            QP_TRY(b_solve(R, s, slot));
        }
        if (p.longitudinal_b) QP_TRY(bx_solve(R, s, slot));
        if (w->env)
            MLAUNCH(k_mx_chi_store, dim3(R.gb), 256, 0, (const double *)w->chi, w->env->chi3, w->ctl, slot, R.nn);
    }
    MLAUNCH(k_mx_out, dim3(R.gb, R.nw), 256, 0, w->ex, w->ey, w->ez, (const double *)w->psi_b[0],
            (const double *)w->psi_b[1], (const double *)w->by, (const double *)w->bz, p.ny, p.nz, p.dx, p.dy, p.dz,
            first ? 0 : 1, w->ctl, slot, 0);
    if (w->bunch)
        QP_TRY(bunch_slice_gather(w->bunch, w->ex, w->ey, w->ez, w->by, w->bz, (const Ctl *)w->ctl, slot, 0, s));
    return QP_OK;
}

int enqueue_slice(MxRun &R, hipStream_t s, int i) {
    return with_tile(R.pt, R.cap, [&](auto P, auto C) { return enqueue_slice_t<decltype(P)::value, decltype(C)::value>(R, s, i); });
}

int resort(MxRun &R, hipStream_t s) {
    Wake *w = R.w;
    MixedState *m = R.m;
    const int ntk = (w->p.nz + kTile - 1) / kTile, ntj = (w->p.ny + kTile - 1) / kTile;
    const int nbins = ntk * ntj * kTileBins, nchunks = (nbins + kScanChunk - 1) / kScanChunk;
    if (nchunks > kScanChunk) {   // the totals are scanned by one block
        set_error("re-binning: grid too large for the bin scan");
        return QP_ERR_ARG;
    }
    const dim3 pg(m->pstride / kPT, 1);
    QP_HIP_CHECK(hipMemsetAsync(m->count, 0, sizeof(unsigned) * nbins, s));
    MLAUNCH(k_sort_hist, pg, kPT, 0, m->P, m->count, ntk);
    MLAUNCH(k_scan_chunks, dim3(nchunks), kScanTB, 0, m->count, m->cursor, m->totals, nbins);
    MLAUNCH(k_scan_chunks, dim3(1), kScanTB, 0, m->totals, (unsigned *)nullptr, m->totals + nchunks, nchunks);
    MLAUNCH(k_scan_add, dim3((nbins + kPT - 1) / kPT), kPT, 0, m->count, (const unsigned *)m->totals, nbins);
    MLAUNCH(k_sort_scatter, pg, kPT, 0, m->P, m->Q, (const unsigned *)m->count, m->cursor, (const int *)m->perm,
            m->perm2, ntk);
    MLAUNCH(k_sort_copyback, pg, kPT, 0, m->P, m->Q, m->perm, (const int *)m->perm2);
    return QP_OK;
}

// Stage the planes of one member (offset go) and call the batch callback.
int deliver_member(Wake *w, batch_cb cb, void *user, int member, int slice, int kind, const double *psi,
                   const double *rho, const double *jx, const double *jy, const double *jz, const double *ex,
                   const double *ey, const double *ez, const double *by, const double *bz, const double *chi,
                   const double *bx = nullptr) {
    const size_t go = (size_t)member * w->nn();
    auto off = [&](const double *a) -> const double * { return a ? a + go : nullptr; };
    struct Fwd {
        batch_cb cb;
        void *user;
        int member;
    } f{cb, user, member};
    auto tramp = [](const qp_slice_view *v, void *u) -> int {
        const Fwd *ff = static_cast<const Fwd *>(u);
        return ff->cb(ff->member, v, ff->user);
    };
    return deliver_slice(w, tramp, &f, slice, kind, off(psi), off(rho), off(jx), off(jy), off(jz), off(ex),
                         off(ey), off(ez), off(by), off(bz), off(chi), off(bx));
}

constexpr int kHostSlots = 4;   // pinned host slots of the streamed output
constexpr int kSnapPlanes = 12; // psi, rho, jx, jy, jz, ex, ey, ez, by, bz, chi, bx

// Slot layout: plane f of member m at (f * nw + m) * nn, then the nw control
// blocks (rounded up to whole doubles).
int ensure_snapshots(Wake *w, MixedState *m) {
    const size_t nn = w->nn(), nw = (size_t)w->nw;
    const size_t ctl_words = (sizeof(Ctl) * nw + sizeof(double) - 1) / sizeof(double);
    const size_t words = kSnapPlanes * nw * nn + ctl_words;
    if (m->snap_dev && m->snap_words == words) return QP_OK;
    if (m->snap_dev) (void)hipFree(m->snap_dev);
    if (m->snap_host) (void)hipHostFree(m->snap_host);
    m->snap_dev = nullptr;
    m->snap_host = nullptr;
    QP_HIP_CHECK(hipMalloc(&m->snap_dev, sizeof(double) * 2 * words));
    QP_HIP_CHECK(hipHostMalloc(&m->snap_host, sizeof(double) * kHostSlots * words, hipHostMallocDefault));
    m->snap_words = words;
    if (!m->copy_stream) QP_HIP_CHECK(hipStreamCreateWithFlags(&m->copy_stream, hipStreamNonBlocking));
    for (hipEvent_t *e : {&m->ev_done[0], &m->ev_done[1], &m->ev_copied[0], &m->ev_copied[1]})
        if (!*e) QP_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    while ((int)m->ev_host.size() < kHostSlots) {
        hipEvent_t e;
        QP_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        m->ev_host.push_back(e);
    }
    return QP_OK;
}

// Snapshot of slice i's output planes and control blocks into device slot d.
int snapshot(Wake *w, MixedState *m, hipStream_t s, int slot, int d) {
    const size_t nn = w->nn(), nw = (size_t)w->nw, plane = nn * nw;
    double *dst = m->snap_dev + (size_t)d * m->snap_words;
    const double *src[kSnapPlanes] = {w->psi_b[slot], w->rho, w->jx, m->jbuf[1 - slot][0], m->jbuf[1 - slot][1],
                                      w->ex, w->ey, w->ez, w->by, w->bz, w->chi,
                                      w->p.longitudinal_b ? w->bx : nullptr};
    for (int f = 0; f < kSnapPlanes; ++f)
        if (src[f])
            QP_HIP_CHECK(hipMemcpyAsync(dst + f * plane, src[f], sizeof(double) * plane, hipMemcpyDeviceToDevice, s));
    QP_HIP_CHECK(hipMemcpyAsync(dst + kSnapPlanes * plane, w->ctl, sizeof(Ctl) * nw, hipMemcpyDeviceToDevice, s));
    return QP_OK;
}

// View of member mb in a host slot.
qp_slice_view host_view(const Wake *w, const double *slot, int mb, int slice, int kind) {
    const size_t nn = w->nn(), nw = (size_t)w->nw;
    auto pl = [&](int f) { return slot + ((size_t)f * nw + mb) * nn; };
    qp_slice_view v;
    v.slice = slice;
    v.kind = kind;
    v.psi = pl(0);
    v.rho = pl(1);
    v.jx = pl(2);
    v.jy = pl(3);
    v.jz = pl(4);
    v.ex = pl(5);
    v.ey = pl(6);
    v.ez = pl(7);
    v.by = pl(8);
    v.bz = pl(9);
    v.chi = pl(10);
    v.bx = w->p.longitudinal_b ? pl(11) : nullptr;
    return v;
}

// The writer side of the streamed output: jobs (slice, host slot) in slice
// order; slots are handed back as they are consumed.
struct OutQueue {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<int, int>> jobs;  // (slice, host slot)
    std::vector<char> slot_free;
    bool closed = false;
    std::atomic<bool> all_stopped{false};
    std::atomic<bool> failed{false};
    std::string error;
};

void writer_loop(Wake *w, MixedState *m, OutQueue *q, batch_cb cb, void *user, std::vector<char> *delivered_stop) {
    (void)hipSetDevice(w->ctx->device);
    const int nw = w->nw;
    const size_t plane = w->nn() * nw;
    const std::vector<double> zero(w->nn(), 0.0);  // the planes a stopped slice does not have
    for (;;) {
        std::pair<int, int> job;
        {
            std::unique_lock<std::mutex> lk(q->mu);
            q->cv.wait(lk, [&] { return q->closed || !q->jobs.empty(); });
            if (q->jobs.empty()) return;
            job = q->jobs.front();
            q->jobs.pop_front();
        }
        const int h = job.second;
        if (!q->failed && hipEventSynchronize(m->ev_host[h]) != hipSuccess) {
            q->error = "streamed output: copy of a slice failed";
            q->failed = true;
        }
        if (!q->failed) {
            const double *slot = m->snap_host + (size_t)h * m->snap_words;
            const Ctl *hc = reinterpret_cast<const Ctl *>(slot + kSnapPlanes * plane);
            int running = 0;
            for (int mb = 0; mb < nw && !q->failed; ++mb) {
                if ((*delivered_stop)[mb]) continue;
                qp_slice_view v;
                if (hc[mb].stop_acc) {
                    v = host_view(w, slot, mb, hc[mb].stop_slice, QP_SLICE_STOPPED);
                    v.psi = v.rho = v.jx = v.jy = v.jz = v.ex = v.ey = v.ez = zero.data();
                    v.bx = nullptr;
                    (*delivered_stop)[mb] = 1;
                } else {
                    ++running;
                    v = host_view(w, slot, mb, job.first, QP_SLICE_NORMAL);
                }
                const int r = cb(mb, &v, user);
                if (r != 0) {
                    q->error = "slice callback returned " + std::to_string(r);
                    q->failed = true;
                }
            }
            if (!running) q->all_stopped = true;
        }
        {
            std::lock_guard<std::mutex> lk(q->mu);
            q->slot_free[h] = 1;
        }
        q->cv.notify_all();
    }
}

}  // namespace

// Pass set-up of one wake or batch: device initial state, launch geometry,
// lasers, zeroed fields and control blocks (src/System_3d.cpp:262-316).
int pass_setup(Wake *w, MxRun &R) {
    const qp_wake_params &p = w->p;
    hipStream_t s = w->stream();
    QP_TRY(wake_upload_mixed(w));
    MixedState *m = mstate(w);
    const int nw = w->nw;
    R.w = w;
    R.m = m;
    R.nn = w->nn();
    R.nw = nw;
    R.pt = w->tile_pt;
    R.cap = w->tile_cap;
    R.pb = m->pstride / R.pt;
    R.gb = (int)((R.nn + 255) / 256);
    R.sstride = (size_t)w->S.n1 * w->S.nh;
    R.g = MxGeom{p.ny, p.nz, p.dy, p.dz, p.ly, p.lz, p.dx, 1.0 / p.dy, 1.0 / p.dz};
    if (w->bunch && w->bunch->epoch != w->bunch_epoch) {  // the bunch's particle arrays moved
        w->bunch_epoch = w->bunch->epoch;
        ++w->inputs_epoch;
    }
    for (int sl = 0; sl < 2; ++sl)
        R.si[sl] = MxSliceIn{w->asq3(), w->asq2, w->rho_b_src(), w->jx_b_src(), w->bsrc_stride(), sl};
    // captured slice steps hold the driver pointers and flags of their inputs
    if (m->graph_epoch != w->inputs_epoch || m->graph_pt != R.pt || m->graph_cap != R.cap) {
        for (auto &g : m->gexec)
            if (g) (void)hipGraphExecDestroy(g);
        m->gexec[0] = m->gexec[1] = nullptr;
        m->graph_epoch = w->inputs_epoch;
        m->graph_pt = R.pt;
        m->graph_cap = R.cap;
    }
    R.ldsS = sizeof(double) * 3 * R.cap;

    // members' lasers (a member without one gets amplitude 0)
    std::vector<LaserGauss> hl(nw);
    for (int i = 0; i < nw; ++i)
        hl[i] = w->mem[i].laser_set ? w->mem[i].laser : LaserGauss{0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0};
    QP_HIP_CHECK(hipMemcpyAsync(w->d_laser, hl.data(), sizeof(LaserGauss) * nw, hipMemcpyHostToDevice, s));
    // zero the fields (src/System_3d.cpp:265-316) and the control blocks
    const size_t plane = sizeof(double) * R.nn * nw;
    double *zero_list[] = {w->psi_b[0], w->psi_b[1], w->psim, w->by, w->bz, w->rho_ion, w->src, w->rho, w->jx,
                           w->chi, w->ex, w->ey, w->ez, w->asq2, w->bx};
    for (double *a : zero_list) QP_HIP_CHECK(hipMemsetAsync(a, 0, plane, s));
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) QP_HIP_CHECK(hipMemsetAsync(m->jbuf[a][b], 0, plane, s));
    std::vector<Ctl> hc0(nw);
    for (Ctl &c : hc0) {
        c = Ctl{};
        c.slice_of[(p.nx - 1) & 1] = p.nx - 1;
        c.slice_of[p.nx & 1] = p.nx;
    }
    QP_HIP_CHECK(hipMemcpyAsync(w->ctl, hc0.data(), sizeof(Ctl) * nw, hipMemcpyHostToDevice, s));
    return QP_OK;
}

// First launches of a pass: rho_ion (src/System_3d.cpp:318-323) and the bunch's
// field buffers.
int pass_start(Wake *w, MxRun &R) {
    hipStream_t s = w->stream();
    MLAUNCH(k_mx_ion, dim3(R.m->pstride / kPT, R.nw), kPT, 0, R.m->P, R.g, w->rho_ion);
    if (w->bunch) QP_TRY(bunch_clear_fields(w->bunch, s));
    return QP_OK;
}

// Inline per-slice output of slice i (after its step): every member still
// running gets output_step(i); a member whose stop rule fired gets its stopped
// slice once.  `running` counts the members still running.
int deliver_inline(Wake *w, MxRun &R, batch_cb cb, void *user, int i, std::vector<char> &delivered_stop,
                   std::vector<Ctl> &hc, int &running) {
    const qp_wake_params &p = w->p;
    MixedState *m = R.m;
    const int nw = R.nw, slot = i & 1;
    QP_HIP_CHECK(hipStreamSynchronize(w->stream()));
    QP_HIP_CHECK(copy_wait(hc.data(), w->ctl, sizeof(Ctl) * nw, hipMemcpyDeviceToHost, w->stream()));
    running = 0;
    for (int mb = 0; mb < nw; ++mb) {
        if (delivered_stop[mb]) continue;
        if (hc[mb].stop_acc) {
            QP_TRY(deliver_member(w, cb, user, mb, hc[mb].stop_slice, QP_SLICE_STOPPED, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, nullptr, w->by, w->bz, w->chi));
            delivered_stop[mb] = 1;
            continue;
        }
        ++running;
        QP_TRY(deliver_member(w, cb, user, mb, i, QP_SLICE_NORMAL, w->psi_b[slot], w->rho, w->jx,
                              m->jbuf[1 - slot][0], m->jbuf[1 - slot][1], w->ex, w->ey, w->ez, w->by, w->bz, w->chi,
                              p.longitudinal_b ? w->bx : nullptr));
    }
    return QP_OK;
}

// Launches after the last slice of a whole pass: output_step(0), the bunch's
// slice-0 gather and an enqueued envelope step.
int pass_final_launch(Wake *w, MxRun &R) {
    const qp_wake_params &p = w->p;
    hipStream_t s = w->stream();
    const int nw = R.nw;
    // output_step(0) after the loop (src/System_3d.cpp:622); the kernel picks each
    // member's psi slot from its own stop slice
    hipLaunchKernelGGL(k_mx_out, dim3(R.gb, nw), dim3(256), 0, s, w->ex, w->ey, w->ez,
                       (const double *)w->psi_b[0], (const double *)w->psi_b[1], (const double *)w->by,
                       (const double *)w->bz, p.ny, p.nz, p.dx, p.dy, p.dz, p.nx > 1 ? 1 : 0, w->ctl, 0, 1);
    QP_HIP_CHECK(hipGetLastError());
    // bunch bucket 0: slice 0's fields are final only after output_step(0)
    if (w->bunch)
        QP_TRY(bunch_slice_gather(w->bunch, w->ex, w->ey, w->ez, w->by, w->bz, (const Ctl *)w->ctl, -1, 0, s));
    // the envelope step after the pass (chi of slices the pass did not reach is
    // the last one stored)
    if (w->env_step()) QP_TRY(envelope_step(w->env, w->env_dtau, w->env_chi_it, w->p.use_graph, s));
    return QP_OK;
}

// After the pass has drained: stop slices, the stopped-slice and final
// (output_step(0)) deliveries.  any_stop gets 1 if a member stopped.
int pass_finish(Wake *w, MxRun &R, batch_cb cb, void *user, int i_end, std::vector<char> &delivered_stop,
                int *stop_slices, int &any_stop) {
    const qp_wake_params &p = w->p;
    MixedState *m = R.m;
    const int nw = R.nw;
    std::vector<Ctl> hc(nw);
    QP_HIP_CHECK(copy_wait(hc.data(), w->ctl, sizeof(Ctl) * nw, hipMemcpyDeviceToHost, w->stream()));
    any_stop = 0;
    w->win_blocks.assign(nw, 0);
    w->win_ovf.assign(nw, 0);
    for (int mb = 0; mb < nw; ++mb) {
        const bool stopped = hc[mb].stop_acc != 0;
        any_stop |= stopped ? 1 : 0;
        // deposit launches of the slices this member ran: A, then S or K x (B, C)
        long long launches = 0;
        for (int i = p.nx - 1; i >= (stopped ? hc[mb].stop_slice : i_end); --i)
            launches += 1 + ((i == 0 || p.K == 0) ? 1 : 2 * p.K);
        w->win_blocks[mb] = launches * (long long)((w->mem[mb].np + R.pt - 1) / R.pt);
        w->win_ovf[mb] = hc[mb].ovf;
        if (stopped && cb && !delivered_stop[mb]) {
            QP_TRY(deliver_member(w, cb, user, mb, hc[mb].stop_slice, QP_SLICE_STOPPED, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, nullptr, w->by, w->bz, w->chi));
        }
        if (cb && i_end == 0) {
            const int cslot = (stopped ? hc[mb].stop_slice : 0) & 1;
            // a stopped member's B and chi planes hold the stopped slice: zero planes
            // instead (zero2 has a zero plane for every member)
            const double *by0 = stopped ? w->zero2 : w->by;
            const double *bz0 = stopped ? w->zero2 : w->bz;
            const double *chi0 = stopped ? w->zero2 : w->chi;
            const double *bx0 = p.longitudinal_b ? (stopped ? w->zero2 : w->bx) : nullptr;
            QP_TRY(deliver_member(w, cb, user, mb, 0, QP_SLICE_FINAL, w->psi_b[cslot], w->rho, w->jx,
                                  m->jbuf[1 - cslot][0], m->jbuf[1 - cslot][1], w->ex, w->ey, w->ez, by0, bz0,
                                  chi0, bx0));
        }
        if (stop_slices) stop_slices[mb] = stopped ? hc[mb].stop_slice : -1;
    }
    return QP_OK;
}

bool mixed_tile_valid(int pt, int cap) { return tile_valid(pt, cap); }
std::string mixed_tiles_json() {
    std::string s = "[";
    for (int i = 0; i < kNumTiles; ++i)
        s += std::string(i ? ", " : "") + "[" + std::to_string(kTiles[i][0]) + ", " + std::to_string(kTiles[i][1]) + "]";
    return s + "]";
}

// ------------------------------------------------------------ tile autotuner
// Picks the particle tile (threads per block, LDS window nodes per grid) of a
// wake or group: each instantiated tile runs the first slices of the pass
// twice (the first run captures the slice step) and the fastest second run
// wins.  Results are cached per (grid sides, members, largest member, K) for
// the process.  QP_TILE=pt,cap forces a tile; a wake with an attached
// envelope keeps the default (a tuning pass would store chi into it).
struct TileKey {
    std::vector<int> k;
    bool operator<(const TileKey &o) const { return k < o.k; }
};
static std::mutex g_tile_mu;
static std::map<TileKey, std::pair<int, int>> g_tile_cache;
constexpr int kTuneSlices = 12;

static bool tile_from_env(int &pt, int &cap) {
    const char *e = getenv("QP_TILE");
    if (!e) return false;
    int a = 0, b = 0;
    if (sscanf(e, "%d,%d", &a, &b) != 2 || !tile_valid(a, b)) return false;
    pt = a;
    cap = b;
    return true;
}

static TileKey tile_key(const std::vector<Wake *> &ws) {
    TileKey key;
    for (Wake *w : ws) {
        size_t mx = 0;
        for (const WakeMember &mb : w->mem) mx = std::max(mx, mb.np);
        key.k.insert(key.k.end(), {w->p.ny, w->p.nz, w->nw, (int)mx, w->p.K});
    }
    return key;
}

// run(pt, cap) runs one short pass with that tile and returns its status;
// last_ms() reads that pass's device time.
template <class Run, class Ms>
int tune_tiles(const TileKey &key, int &pt, int &cap, Run &&run, Ms &&last_ms) {
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        auto it = g_tile_cache.find(key);
        if (it != g_tile_cache.end()) {
            pt = it->second.first;
            cap = it->second.second;
            return QP_OK;
        }
    }
    double best = 1e300;
    int bpt = pt, bcap = cap;
    for (const auto &t : kTiles) {
        for (int rep = 0; rep < 2; ++rep) QP_TRY(run(t[0], t[1]));
        const double ms = last_ms();
        if (ms < best) {
            best = ms;
            bpt = t[0];
            bcap = t[1];
        }
    }
    pt = bpt;
    cap = bcap;
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_cache[key] = std::make_pair(pt, cap);
    return QP_OK;
}

static int tune_wake(Wake *w) {
    w->tile_tuned = true;   // the tuning passes below run with the tile they set
    if (tile_from_env(w->tile_pt, w->tile_cap)) return QP_OK;
    if (w->env || w->p.nx < 3 || w->p.K == 0) return QP_OK;
    const qp_wake_params saved = w->p;
    const int lim = std::min(w->p.nx, kTuneSlices);
    const TileKey key = tile_key({w});
    int pt = w->tile_pt, cap = w->tile_cap;
    int st = tune_tiles(
        key, pt, cap,
        [&](int a, int b) {
            w->tile_pt = a;
            w->tile_cap = b;
            w->p.slice_limit = (saved.slice_limit > 0 && saved.slice_limit < lim) ? saved.slice_limit : lim;
            w->p.async_output = 0;
            const int r = wake_run_mixed(w, nullptr, nullptr, nullptr);
            w->p = saved;
            return r < 0 ? r : QP_OK;
        },
        [&] { return w->last_ms; });
    w->p = saved;
    w->tile_pt = pt;
    w->tile_cap = cap;
    return st;
}

int wake_run_mixed(Wake *w, batch_cb cb, void *user, int *stop_slices) {
    if (!w->tile_tuned && !w->prof_class) QP_TRY(tune_wake(w));
    const qp_wake_params &p = w->p;
    hipStream_t s = w->stream();
    MxRun R;
    QP_TRY(pass_setup(w, R));
    MixedState *m = R.m;
    const int nw = R.nw;
    QP_HIP_CHECK(hipEventRecord(w->ev0, s));
    QP_TRY(pass_start(w, R));

    const bool async = cb && p.async_output && !w->prof_class;
    const bool graph_ok = p.use_graph && (!cb || async) && !w->prof_class;
    std::vector<char> delivered_stop(nw, 0);
    std::vector<Ctl> hc(nw);
    OutQueue q;
    std::thread writer;
    long seq = 0;
    if (async) {
        QP_TRY(ensure_snapshots(w, m));
        q.slot_free.assign(kHostSlots, 1);
        writer = std::thread(writer_loop, w, m, &q, cb, user, &delivered_stop);
    }
    // ends the writer (after the jobs queued so far) on every exit path
    struct Join {
        OutQueue &q;
        std::thread &t;
        ~Join() {
            if (!t.joinable()) return;
            {
                std::lock_guard<std::mutex> lk(q.mu);
                q.closed = true;
            }
            q.cv.notify_all();
            t.join();
        }
    } join{q, writer};
    const int i_end = (p.slice_limit > 0 && p.slice_limit < p.nx) ? p.nx - p.slice_limit : 0;
    for (int i = p.nx - 1; i >= i_end; --i) {
        const int slot = i & 1;
        if (p.resort_every > 0 && i < p.nx - 1 && ((p.nx - 1 - i) % p.resort_every) == 0 && m->P.np)
            QP_TRY(resort(R, s));
        const bool generic = (i < p.nx - 1) && (i > 0) && (p.K > 0);
        if (graph_ok && generic) {
            if (!m->gexec[slot]) {
                hipGraph_t gr;
                QP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int st = enqueue_slice(R, s, i);
                hipError_t e = hipStreamEndCapture(s, &gr);
                if (st < 0) {
                    if (e == hipSuccess) (void)hipGraphDestroy(gr);
                    return st;
                }
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
        if (async && i > 0) {
            // device slot d was last read by the copy of slice i + 2
            const int d = i & 1;
            QP_HIP_CHECK(hipStreamWaitEvent(s, m->ev_copied[d], 0));
            QP_TRY(snapshot(w, m, s, slot, d));
            QP_HIP_CHECK(hipEventRecord(m->ev_done[d], s));
            const int h = (int)(seq++ % kHostSlots);
            {
                std::unique_lock<std::mutex> lk(q.mu);
                q.cv.wait(lk, [&] { return q.slot_free[h] != 0; });
                q.slot_free[h] = 0;
            }
            hipStream_t c = m->copy_stream;
            QP_HIP_CHECK(hipStreamWaitEvent(c, m->ev_done[d], 0));
            QP_HIP_CHECK(hipMemcpyAsync(m->snap_host + (size_t)h * m->snap_words,
                                        m->snap_dev + (size_t)d * m->snap_words, sizeof(double) * m->snap_words,
                                        hipMemcpyDeviceToHost, c));
            QP_HIP_CHECK(hipEventRecord(m->ev_copied[d], c));
            QP_HIP_CHECK(hipEventRecord(m->ev_host[h], c));
            {
                std::lock_guard<std::mutex> lk(q.mu);
                q.jobs.emplace_back(i, h);
            }
            q.cv.notify_all();
            if (q.failed || q.all_stopped) break;
        } else if (cb && i > 0) {
            int running = 0;
            QP_TRY(deliver_inline(w, R, cb, user, i, delivered_stop, hc, running));
            if (!running) break;
        }
    }
    if (async) {  // every streamed slice delivered before the final step
        {
            std::lock_guard<std::mutex> lk(q.mu);
            q.closed = true;
        }
        q.cv.notify_all();
        writer.join();
        if (q.failed) {
            set_error(q.error);
            return QP_ERR_STATE;
        }
    }
    QP_HIP_CHECK(hipStreamSynchronize(s));
    if (i_end == 0) QP_TRY(pass_final_launch(w, R));
    QP_HIP_CHECK(hipEventRecord(w->ev1, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    QP_HIP_CHECK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    w->last_ms = ms;
    int any_stop = 0;
    QP_TRY(pass_finish(w, R, cb, user, i_end, delivered_stop, stop_slices, any_stop));
    return any_stop ? QP_STOPPED : QP_OK;
}


// ------------------------------------------------------------------ groups
namespace {

// Launch tables of a group: per phase and slice parity one argument struct per
// batch, in one device buffer whose layout is fixed by the batch count (the
// captured slice steps hold pointers into it; each run rewrites the contents).
struct GroupTabs {
    GArgsA A[2][kGroupMax];
    GArgsB B[2][kGroupMax];
    GArgsC C[2][kGroupMax];
    GArgsOut O[kGroupMax];
    GArgsRowsF<SrcPsi> PF[2][kGroupMax];
    GArgsCols PC[kGroupMax];
    GArgsRowsI<SinkPsiFinal> PI[2][kGroupMax];   // psi of the slice
    GArgsRowsI<SinkPsiFinal> MI[kGroupMax];      // psi_middle
    GArgsRowsF<SrcB> BF[2][kGroupMax];
    GArgsCols BC[kGroupMax];
    GArgsRowsI<SinkBPair> BI[2][kGroupMax];
};

struct GroupMaps {
    GrpMap part, out;
    GrpMap rows1[2], rows2[2], cols1[2], cols2[2];   // per transform family (sides 2^p, 3 * 2^p)
};

struct GroupState {
    // the second transform family's launches run on this stream, forked from and
    // joined back into the group's (a parallel branch of the captured step)
    hipStream_t side = nullptr;
    std::vector<hipEvent_t> ev;   // fork / join pairs of one slice step
    size_t ev_next = 0;
    GroupTabs *d_tabs = nullptr;
    GroupTabs h_tabs;
    GroupMaps maps{}, graph_maps{};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<unsigned> graph_epochs;
};

int map_total(const GrpMap &m) { return m.start[m.nb]; }

void map_add(GrpMap &m, int b, int gx, int gy) {
    const int i = m.nb;
    m.gx[i] = gx;
    m.bi[i] = b;
    m.start[i + 1] = m.start[i] + gx * gy;
    m.nb = i + 1;
}

// Launch tables and block maps of every batch (R[b] from pass_setup).
void group_tables(Group *G, std::vector<MxRun> &R, GroupTabs &T, GroupMaps &M) {
    M = GroupMaps{};
    for (size_t bi = 0; bi < G->b.size(); ++bi) {
        const int b = (int)bi;
        Wake *w = G->b[b];
        const qp_wake_params &p = w->p;
        MixedState *m = R[b].m;
        const Solver2d &S = w->S;
        const int nw = R[b].nw;
        const int laser = (w->any_laser() && !w->asq3()) ? 1 : 0;
        const double scale = 1.0 / ((double)S.n1 * S.n2);
        for (int sl = 0; sl < 2; ++sl) {
            const MxSliceIn &si = R[b].si[sl];
            T.A[sl][b] = GArgsA{m->P, R[b].g, si, w->psim, w->src, w->asq2, (const LaserGauss *)w->d_laser,
                                laser, w->rho, w->jx, w->chi, w->ctl};
            T.B[sl][b] = GArgsB{m->P, R[b].g, si, w->psi_b[sl], w->by, w->bz, w->src, w->rho, w->jx, w->chi,
                                m->jbuf[sl][0], m->jbuf[sl][1], w->ctl};
            T.C[sl][b] = GArgsC{m->P, R[b].g, si, w->psi_b[sl], w->psim, w->rho, w->jx, w->chi,
                                m->jbuf[sl][0], m->jbuf[sl][1], w->ctl};
            SrcPsi sp{w->src, w->rho_ion, w->rho_b_src(), w->jx_b_src(), w->bsrc_stride(), w->ctl, sl, S.n2, R[b].nn};
            T.PF[sl][b] = GArgsRowsF<SrcPsi>{sp, S.spec[0], S.spec[0], S.n1, S.n2, S.ax2.dev.tw, w->ctl,
                                             R[b].sstride};
            SinkPsiFinal kp{w->psi_b[sl], m->bpart, m->ncb, S.n1, S.n2, p.psi_threshold - 1, 0.0, m->bstride, scale};
            T.PI[sl][b] = GArgsRowsI<SinkPsiFinal>{kp, S.spec[0], S.spec[0], S.n1, S.n2, S.ax2.dev.tw, w->ctl,
                                                   R[b].sstride};
            double *jy = m->jbuf[1 - sl][0], *jz = m->jbuf[1 - sl][1];
            double *jyn = m->jbuf[sl][0], *jzn = m->jbuf[sl][1];
            SrcB sb{w->by, w->bz, w->jx, jy, jz, jyn, jzn, S.n1, S.n2, p.D, 1.0 / p.dx, 1.0 / p.dy, 1.0 / p.dz};
            T.BF[sl][b] = GArgsRowsF<SrcB>{sb, S.spec[0], S.spec[1], S.n1, S.n2, S.ax2.dev.tw, w->ctl, R[b].sstride};
            SinkBPair kb{w->by, w->bz, S.n1, S.n2, p.b_threshold, w->ctl, sl, scale};
            T.BI[sl][b] = GArgsRowsI<SinkBPair>{kb, S.spec[0], S.spec[1], S.n1, S.n2, S.ax2.dev.tw, w->ctl,
                                                R[b].sstride};
        }
        T.O[b] = GArgsOut{w->ex, w->ey, w->ez, w->psi_b[0], w->psi_b[1], w->by, w->bz, p.ny, p.nz, p.dx, p.dy, p.dz,
                          w->ctl};
        T.PC[b] = GArgsCols{S.spec[0], S.spec[0], S.n1, S.nh, S.n2, S.ax1.dev.tw, S.d_lam1, S.d_lam2, 0.0, m->bpart,
                            w->ctl, R[b].sstride, m->bstride};
        T.BC[b] = GArgsCols{S.spec[0], S.spec[1], S.n1, S.nh, S.n2, S.ax1.dev.tw, S.d_lam1, S.d_lam2, p.D, nullptr,
                            w->ctl, R[b].sstride, 0};
        SinkPsiFinal km{w->psim, m->bpart, m->ncb, S.n1, S.n2, p.psi_threshold - 1, 0.0, m->bstride, scale};
        T.MI[b] = GArgsRowsI<SinkPsiFinal>{km, S.spec[0], S.spec[0], S.n1, S.n2, S.ax2.dev.tw, w->ctl,
                                           R[b].sstride};
        // block maps: the launch grids of the single-batch path with 64-thread solver blocks
        const int T2 = kGrpTB * grp_elems(S.n2) / S.n2, T1 = kGrpTB * grp_elems(S.n1) / S.n1;
        const int f2 = S.n2 % 3 == 0 ? 1 : 0, f1 = S.n1 % 3 == 0 ? 1 : 0;   // transform families
        map_add(M.part, b, R[b].pb, nw);
        map_add(M.rows1[f2], b, wrows_blocks_t(S.n1, T2), nw);
        map_add(M.rows2[f2], b, wrows_blocks_t(S.n1, T2), 2 * nw);
        map_add(M.cols1[f1], b, wcols_blocks_t(S.nh, T1), nw);
        map_add(M.cols2[f1], b, wcols_blocks_t(S.nh, T1), 2 * nw);
        map_add(M.out, b, R[b].gb, nw);
    }
}

#define GLAUNCH(kern, map, block, lds, ...)                                                    \
    do {                                                                                       \
        if ((map).nb > 0) {                                                                    \
            hipLaunchKernelGGL(kern, dim3(map_total(map)), dim3(block), lds, s, map, __VA_ARGS__); \
            QP_HIP_CHECK(hipGetLastError());                                                   \
        }                                                                                      \
    } while (0)
// One solver pass of every batch: one launch per transform family present
// (GSOLVE_T for the kernels templated on their source / sink type).
// The two transform families of one solver pass are independent launches with
// few blocks each (latency-bound): with both present and Group::fork_families
// set, the 3 * 2^p family's runs on the side stream beside the 2^p family's.
// The 64-wake group of five grid sizes: 2.50 -> 2.40 ms per slice.  With the
// same wakes in three chains on their own streams (one chain a group of three
// batches) the extra stream was slower (1.99 -> 2.32 ms per slice), so callers
// running several groups or batches at once switch it off.
template <class L0, class L1>
int fam_pair(GroupState *st, hipStream_t s, bool both, L0 &&l0, L1 &&l1) {
    if (!both) {
        QP_TRY(l0(s));
        return l1(s);
    }
    if (st->ev_next + 2 > st->ev.size()) {
        set_error("group: fork events exhausted");
        return QP_ERR_HIP;
    }
    hipEvent_t ef = st->ev[st->ev_next++], ej = st->ev[st->ev_next++];
    QP_HIP_CHECK(hipEventRecord(ef, s));
    QP_HIP_CHECK(hipStreamWaitEvent(st->side, ef, 0));
    QP_TRY(l1(st->side));
    QP_HIP_CHECK(hipEventRecord(ej, st->side));
    QP_TRY(l0(s));
    QP_HIP_CHECK(hipStreamWaitEvent(s, ej, 0));
    return QP_OK;
}
#define GSOLVE(kern, maps, nspec, tab)                                                                     \
    QP_TRY(fam_pair(                                                                                       \
        st, s, G->fork_families && (maps)[0].nb > 0 && (maps)[1].nb > 0,                                   \
        [&](hipStream_t s) {                                                                               \
            GLAUNCH((kern<0>), maps[0], kGrpTB, sizeof(double2) * 2 * grp_lds(0), tab, nspec, nspec);      \
            return QP_OK;                                                                                  \
        },                                                                                                 \
        [&](hipStream_t s) {                                                                               \
            GLAUNCH((kern<1>), maps[1], kGrpTB, sizeof(double2) * 2 * grp_lds(1), tab, nspec, nspec);      \
            return QP_OK;                                                                                  \
        }))
#define GSOLVE_T(kern, T, maps, nspec, tab)                                                                \
    QP_TRY(fam_pair(                                                                                       \
        st, s, G->fork_families && (maps)[0].nb > 0 && (maps)[1].nb > 0,                                   \
        [&](hipStream_t s) {                                                                               \
            GLAUNCH((kern<T, 0>), maps[0], kGrpTB, sizeof(double2) * 2 * grp_lds(0), tab, nspec, nspec);   \
            return QP_OK;                                                                                  \
        },                                                                                                 \
        [&](hipStream_t s) {                                                                               \
            GLAUNCH((kern<T, 1>), maps[1], kGrpTB, sizeof(double2) * 2 * grp_lds(1), tab, nspec, nspec);   \
            return QP_OK;                                                                                  \
        }))

// One slice step of every batch: enqueue_slice with one launch per phase.
template <int PT, int CAP>
int enqueue_group_t(Group *G, GroupState *st, hipStream_t s, int i) {
    const qp_wake_params &p = G->b[0]->p;
    const GroupMaps &M = st->maps;
    GroupTabs *T = st->d_tabs;
    st->ev_next = 0;
    const int slot = i & 1;
    const bool first = (i == p.nx - 1);
    const bool s_phase = (i == 0) || (p.K == 0);
    GLAUNCH((kg_mx_A<PT, CAP>), M.part, PT, 0, T->A[slot], s_phase ? 1 : 0, first ? 0 : (p.K == 0 ? 2 : 1));
    GSOLVE_T(kg_rows_fwd, SrcPsi, M.rows1, 1, T->PF[slot]);
    GSOLVE(kg_cols, M.cols1, 1, T->PC);
    GSOLVE_T(kg_rows_inv, SinkPsiFinal, M.rows1, 1, T->PI[slot]);
    if (s_phase) {
        GLAUNCH((kg_mx_S<PT, CAP>), M.part, PT, sizeof(double) * 3 * CAP, T->C[slot]);
        if (i == 0) return QP_OK;
    } else {
        for (int it = 0; it < p.K; ++it) {
            const bool first_it = it == 0, last_it = it == p.K - 1;
            GLAUNCH((kg_mx_B<PT, CAP>), M.part, PT, 0, T->B[slot], first_it ? 1 : 0, last_it ? 1 : 0);
            GSOLVE_T(kg_rows_fwd, SrcPsi, M.rows1, 1, T->PF[slot]);
            GSOLVE(kg_cols, M.cols1, 1, T->PC);
            GSOLVE_T(kg_rows_inv, SinkPsiFinal, M.rows1, 1, T->MI);
            const size_t ldsc = sizeof(double) * CAP * (3 + (first_it ? 1 : 0) + (last_it ? 1 : 0));
            if (first_it && last_it)
                GLAUNCH((kg_mx_C<PT, CAP, true, true>), M.part, PT, ldsc, T->C[slot]);
            else if (first_it)
                GLAUNCH((kg_mx_C<PT, CAP, false, true>), M.part, PT, ldsc, T->C[slot]);
            else if (last_it)
                GLAUNCH((kg_mx_C<PT, CAP, true, false>), M.part, PT, ldsc, T->C[slot]);
            else
                GLAUNCH((kg_mx_C<PT, CAP, false, false>), M.part, PT, ldsc, T->C[slot]);
            GSOLVE_T(kg_rows_fwd, SrcB, M.rows2, 2, T->BF[slot]);
            GSOLVE(kg_cols, M.cols2, 2, T->BC);
            GSOLVE_T(kg_rows_inv, SinkBPair, M.rows2, 2, T->BI[slot]);
        }
    }
    GLAUNCH(kg_mx_out, M.out, 256, 0, T->O, first ? 0 : 1, slot, 0);
    return QP_OK;
}

int enqueue_group(Group *G, GroupState *st, hipStream_t s, int i) {
    return with_tile(G->b[0]->tile_pt, G->b[0]->tile_cap, [&](auto P, auto C) {
        return enqueue_group_t<decltype(P)::value, decltype(C)::value>(G, st, s, i);
    });
}

// What a group launch takes: mixed batches on one context with the same slice
// count, iterations and pass length, Gaussian lasers only, inline output, and
// grid sides the group's solver kernels carry.
int group_check(Group *G) {
    if (G->b.empty() || (int)G->b.size() > kGroupMax) {
        set_error("group: 1 to " + std::to_string(kGroupMax) + " batches");
        return QP_ERR_ARG;
    }
    const qp_wake_params &p0 = G->b[0]->p;
    for (Wake *w : G->b) {
        const qp_wake_params &p = w->p;
        if (w->ctx != G->ctx || p.precision != QP_PREC_MIXED) {
            set_error("group: every batch must be a mixed-precision handle of the group's context");
            return QP_ERR_ARG;
        }
        if (p.nx != p0.nx || p.K != p0.K || p.slice_limit != p0.slice_limit) {
            set_error("group: batches must share nx, K and slice_limit (they advance in lockstep)");
            return QP_ERR_ARG;
        }
        if (p.async_output || p.resort_every || p.longitudinal_b) {
            set_error("group: async_output, resort_every and longitudinal_b are single-handle options");
            return QP_ERR_ARG;
        }
        if (w->env || w->bunch || w->a_sqr3 || w->rho_b3 || w->jx_b3) {
            set_error("group: batches take Gaussian lasers only (no envelope, bunch, 3D a_sqr or bunch sources)");
            return QP_ERR_STATE;
        }
        if (!w->S.wave || !grp_side_ok(p.ny) || !grp_side_ok(p.nz)) {
            set_error("group: grid sides must be 2^p in [16, 256] or 48, 96, 192 (got " + std::to_string(p.ny) + " x " +
                      std::to_string(p.nz) + ")");
            return QP_ERR_ARG;
        }
    }
    return QP_OK;
}

// Adapts the group callback to one batch's member callback.
struct GroupCbCtx {
    group_cb cb;
    void *user;
    int batch;
};
int group_member_cb(int member, const qp_slice_view *v, void *u) {
    const GroupCbCtx *c = static_cast<const GroupCbCtx *>(u);
    return c->cb(c->batch, member, v, c->user);
}

}  // namespace

void group_destroy_mixed(Group *G) {
    GroupState *st = static_cast<GroupState *>(G->dev);
    if (!st) return;
    for (auto &g : st->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (st->d_tabs) (void)hipFree(st->d_tabs);
    for (hipEvent_t e : st->ev)
        if (e) (void)hipEventDestroy(e);
    if (st->side) (void)hipStreamDestroy(st->side);
    delete st;
    G->dev = nullptr;
}

// Tile autotuner of a group (see tune_tiles): one tile for all its launches.
static int tune_group(Group *G) {
    G->tile_tuned = true;
    int pt = G->tile_pt, cap = G->tile_cap;
    if (!tile_from_env(pt, cap) && G->b[0]->p.nx >= 3 && G->b[0]->p.K > 0) {
        std::vector<qp_wake_params> saved;
        for (Wake *w : G->b) saved.push_back(w->p);
        const int lim = std::min(G->b[0]->p.nx, kTuneSlices);
        QP_TRY(tune_tiles(
            tile_key(G->b), pt, cap,
            [&](int a, int b) {
                G->tile_pt = a;
                G->tile_cap = b;
                for (Wake *w : G->b)
                    w->p.slice_limit = (w->p.slice_limit > 0 && w->p.slice_limit < lim) ? w->p.slice_limit : lim;
                const int r = group_run_mixed(G, nullptr, nullptr, nullptr);
                for (size_t i = 0; i < G->b.size(); ++i) G->b[i]->p = saved[i];
                return r < 0 ? r : QP_OK;
            },
            [&] { return G->last_ms; }));
    }
    G->tile_pt = pt;
    G->tile_cap = cap;
    return QP_OK;
}

int group_run_mixed(Group *G, group_cb cb, void *user, int *stop_slices) {
    QP_TRY(group_check(G));
    if (!G->tile_tuned) QP_TRY(tune_group(G));
    for (Wake *w : G->b) {   // every batch launches with the group's tile
        w->tile_pt = G->tile_pt;
        w->tile_cap = G->tile_cap;
        w->tile_tuned = true;
    }
    const int nb = (int)G->b.size();
    hipStream_t s = G->ctx->stream;
    GroupState *st = static_cast<GroupState *>(G->dev);
    if (!st) {
        st = new GroupState();
        G->dev = st;
        QP_HIP_CHECK(hipMalloc(&st->d_tabs, sizeof(GroupTabs)));
        QP_HIP_CHECK(hipStreamCreateWithFlags(&st->side, hipStreamNonBlocking));
    }
    // fork / join events: two per solver pass of one slice step (3 + 6 K passes)
    while (st->ev.size() < (size_t)(2 * (3 + 6 * G->b[0]->p.K))) {
        hipEvent_t e = nullptr;
        QP_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        st->ev.push_back(e);
    }
    std::vector<MxRun> R(nb);
    for (int b = 0; b < nb; ++b) QP_TRY(pass_setup(G->b[b], R[b]));
    group_tables(G, R, st->h_tabs, st->maps);
    QP_HIP_CHECK(hipMemcpyAsync(st->d_tabs, &st->h_tabs, sizeof(GroupTabs), hipMemcpyHostToDevice, s));
    // captured steps hold the block maps and the batches' driver flags
    std::vector<unsigned> epochs(nb);
    for (int b = 0; b < nb; ++b) epochs[b] = G->b[b]->inputs_epoch;
    epochs.push_back((unsigned)G->tile_cap);   // the maps carry the block counts of the tile, not its capacity
    epochs.push_back(G->fork_families ? 1u : 0u);
    if (std::memcmp(&st->maps, &st->graph_maps, sizeof(GroupMaps)) != 0 || epochs != st->graph_epochs) {
        for (auto &g : st->gexec)
            if (g) (void)hipGraphExecDestroy(g);
        st->gexec[0] = st->gexec[1] = nullptr;
        st->graph_maps = st->maps;
        st->graph_epochs = epochs;
    }
    QP_HIP_CHECK(hipEventRecord(G->ev0, s));
    for (int b = 0; b < nb; ++b) QP_TRY(pass_start(G->b[b], R[b]));

    const qp_wake_params &p = G->b[0]->p;
    const bool graph_ok = p.use_graph && !cb;
    std::vector<std::vector<char>> delivered_stop(nb);
    std::vector<std::vector<Ctl>> hc(nb);
    std::vector<GroupCbCtx> cctx(nb);
    for (int b = 0; b < nb; ++b) {
        delivered_stop[b].assign(R[b].nw, 0);
        hc[b].resize(R[b].nw);
        cctx[b] = GroupCbCtx{cb, user, b};
    }
    const int i_end = (p.slice_limit > 0 && p.slice_limit < p.nx) ? p.nx - p.slice_limit : 0;
    for (int i = p.nx - 1; i >= i_end; --i) {
        const int slot = i & 1;
        const bool generic = (i < p.nx - 1) && (i > 0) && (p.K > 0);
        if (graph_ok && generic) {
            if (!st->gexec[slot]) {
                hipGraph_t gr;
                QP_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int rc = enqueue_group(G, st, s, i);
                hipError_t e = hipStreamEndCapture(s, &gr);
                if (rc < 0) {
                    if (e == hipSuccess) (void)hipGraphDestroy(gr);
                    return rc;
                }
                if (e != hipSuccess) {
                    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                    return QP_ERR_HIP;
                }
                e = hipGraphInstantiate(&st->gexec[slot], gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) {
                    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                    return QP_ERR_HIP;
                }
            }
            QP_HIP_CHECK(hipGraphLaunch(st->gexec[slot], s));
        } else {
            QP_TRY(enqueue_group(G, st, s, i));
        }
        if (cb && i > 0) {
            int running = 0;
            for (int b = 0; b < nb; ++b) {
                int rb = 0;
                QP_TRY(deliver_inline(G->b[b], R[b], group_member_cb, &cctx[b], i, delivered_stop[b], hc[b], rb));
                running += rb;
            }
            if (!running) break;
        }
    }
    QP_HIP_CHECK(hipStreamSynchronize(s));
    if (i_end == 0)
        for (int b = 0; b < nb; ++b) QP_TRY(pass_final_launch(G->b[b], R[b]));
    QP_HIP_CHECK(hipEventRecord(G->ev1, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    QP_HIP_CHECK(hipEventElapsedTime(&ms, G->ev0, G->ev1));
    G->last_ms = ms;
    int any = 0, off = 0;
    for (int b = 0; b < nb; ++b) {
        int a = 0;
        QP_TRY(pass_finish(G->b[b], R[b], cb ? group_member_cb : nullptr, &cctx[b], i_end, delivered_stop[b],
                           stop_slices ? stop_slices + off : nullptr, a));
        any |= a;
        off += R[b].nw;
    }
    return any ? QP_STOPPED : QP_OK;
}

}  // namespace qp
