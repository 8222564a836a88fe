// qp_mixed_kernels.h -- device code of the fused mixed-precision slice step.
//
// Particle layout (structure of arrays, 36 B persistent + 16 B scratch):
//   cell  uint32   (j << 16) | k          node-origin cell of the particle
//   off   uint2    in-cell offsets (y, z) in units of 2^-32 cell (Y = j + off.x 2^-32)
//   p     double2  transverse momenta (y, z) (they accumulate over every slice)
//   w     double   macro-particle weight (wake_particle_2d::n); in single precision
//                  its rounding alone (up to 6e-8 relative for a density ramp) moved
//                  e_x of a 256^2 x 800-slice wake at a0 = 2.3 by 1.1e-5 of its
//                  maximum against the FP64 oracle
//   ps    double2  the current iteration's predicted momenta p* (scratch).  As
//                  an FP32 increment p* - p its rounding alone moved e_x of a
//                  256^2 x 800-slice wake at a0 = 2.3 by 1.1e-5 of its maximum
//                  against the FP64 oracle (the rounding enters every commit)
// Arithmetic is FP64 throughout (without packed FP32, FP32 and FP64 add, mul
// and FMA issue at the same rate on gfx950); only the position is narrowed.
// Positions are advanced in FP64 cell units and re-split with round-to-nearest
// into cell + offset, so the stored position keeps ~2^-32 cell resolution
// over thousands of slices.
//
// The 35 particle sweeps of a reference slice (src/System_3d.cpp:328-619)
// become 2K+1 launches: A (commit + advance + psi source deposit), then per
// iteration B (gamma, predictor, half step, psi_middle source deposit) and
// C (transverse current at the half step, gamma/px, jx deposits; chi in the
// first iteration of a slice, rho in the last).
#pragma once

#include "qp_solver_kernels.h"
#include "qp_wake.h"

namespace qp {

// Particle tiles: PT threads (= particles) per block and CAP LDS window nodes
// per deposited grid are template parameters of the particle kernels, picked
// per wake by the tile autotuner among the instantiated pairs (qp_wake_mixed.hip);
// kPT / kWinCap are the defaults, and kPT is the block of the re-binning and
// rho_ion kernels.
constexpr int kPT = 256;
constexpr int kWinCap = 640;
constexpr double kTwo32 = 4294967296.0;
constexpr double kInvTwo32 = 1.0 / 4294967296.0;

struct MxGeom {
    int n1, n2;
    double d1, d2, ly, lz, dx;
    double inv_d1, inv_d2;  // 1/dy, 1/dz: the fused path multiplies instead of dividing
};

// Member m of a batch owns slots [m * pstride, m * pstride + npw[m]) of every
// array (npw null: one member with np particles).
// Pairs that every kernel reads together sit in one element, so a particle's
// state is five loads: cell, offsets (y, z), momenta (y, z), weight and the
// p* increments (y, z).
struct MxPart {
    uint32_t *cell;
    uint2 *off;    // in-cell offsets (y, z) in units of 2^-32 of a cell
    double2 *p;    // p_y, p_z
    double *w;
    double2 *ps;   // p* (y, z)
    int np;
    int pstride;
    const int *npw;
};

__device__ __forceinline__ int mx_np(const MxPart &P, int wk) { return P.npw ? P.npw[wk] : P.np; }

// Slice-dependent inputs resolved on the device from the slice counter.
struct MxSliceIn {
    const double *a_sqr3;   // nx*n1*n2 or null
    const double *asq2;     // laser plane buffer (used when a_sqr3 is null)
    const double *rho_b3, *jx_b3;  // bunch sources: slice i at + i * bstride (0: one streamed plane)
    size_t bstride;
    int slot;               // which ctl->slice_of[] entry holds this step's slice
};

__device__ __forceinline__ int mx_slice(const Ctl *c, const MxSliceIn &si) { return c->slice_of[si.slot]; }
// <a^2> plane of the block's member (go = member * nn; the 3D array and the
// bunch sources exist for single wakes only).
__device__ __forceinline__ const double *mx_asq(const Ctl *c, const MxSliceIn &si, size_t nn, size_t go) {
    return si.a_sqr3 ? si.a_sqr3 + (size_t)mx_slice(c, si) * nn : si.asq2 + go;
}

// ----------------------------------------------------------- positions
struct Pos {
    int j, k;
    double fy, fz;
};

__device__ __forceinline__ Pos load_pos(const MxPart &P, int p) {
    const uint32_t c = P.cell[p];
    Pos r;
    r.j = (int)(c >> 16);
    r.k = (int)(c & 0xffffu);
    const uint2 o = P.off[p];
    r.fy = (double)o.x * kInvTwo32;
    r.fz = (double)o.y * kInvTwo32;
    return r;
}

// Split a cell-unit coordinate Y in [0, n) (after wrap) into cell and offset,
// offset rounded to the nearest 2^-32.
__device__ __forceinline__ void split_coord(double Y, int n, int &j, uint32_t &o) {
    double fl = floor(Y);
    int jj = (int)fl;
    double u = rint((Y - fl) * kTwo32);
    if (u >= kTwo32) {
        u -= kTwo32;
        jj += 1;
    }
    if (jj >= n) jj -= n;
    if ((unsigned)jj >= (unsigned)n) jj = 0;
    j = jj;
    o = (uint32_t)u;
}

// Periodic wrap in cell units (normalize_coordinates, src/System_3d.cpp:668-684).
// A step moves a particle by less than one box width, so fmod reduces to one
// subtraction (exact: Y - n with n <= Y < 2n is representable) or one
// addition (the reference's y += l); fmod stays for anything further out.
__device__ __forceinline__ double wrap_cells(double Y, int n) {
    const double dn = (double)n;
    if (Y >= 0.0 && Y < dn) return Y;
    if (Y >= dn && Y < 2.0 * dn) {
        Y -= dn;
    } else if (Y < 0.0 && Y >= -dn) {
        Y += dn;
    } else {
        Y = fmod(Y, dn);
        if (Y < 0) Y += dn;
    }
    if (Y >= dn) Y = 0.0;
    return Y;
}

// Element i of a grid through a 32-bit byte offset from the (uniform) base:
// the compiler then uses the scalar-base + 32-bit-offset load form instead of
// 64-bit address arithmetic per lane.  Grids hold at most 1024 x 1024 nodes
// per member, and member bases are applied to the pointer.
__device__ __forceinline__ double ld_at(const double *a, unsigned i) {
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(a) + (size_t)(i << 3));
}
__device__ __forceinline__ double *at(double *a, unsigned i) {
    return reinterpret_cast<double *>(reinterpret_cast<char *>(a) + (size_t)(i << 3));
}
// Node index j * n2 + k with a 24-bit multiply (j, n2 <= 1024).
__device__ __forceinline__ unsigned nidx(int j, int k, int n2) { return __umul24((unsigned)j, (unsigned)n2) + (unsigned)k; }

// Bilinear weights of the four nodes (j1,k1), (j2,k1), (j1,k2), (j2,k2).
__device__ __forceinline__ double bilin(double c11, double c21, double c12, double c22, double yf, double zf) {
    return c11 * (1 - yf) * (1 - zf) + c21 * yf * (1 - zf) + c12 * (1 - yf) * zf + c22 * yf * zf;
}

__device__ __forceinline__ double gather4u(const double *a, int n2, int j1, int j2, int k1, int k2, double yf,
                                           double zf) {
    const unsigned r1 = __umul24((unsigned)j1, (unsigned)n2), r2 = __umul24((unsigned)j2, (unsigned)n2);
    return bilin(ld_at(a, r1 + k1), ld_at(a, r2 + k1), ld_at(a, r1 + k2), ld_at(a, r2 + k2), yf, zf);
}

__device__ __forceinline__ Pos pos_from_cells(double Y, double Z, int n1, int n2) {
    Pos r;
    double fl = floor(Y);
    r.j = (int)fl;
    r.fy = Y - fl;
    if ((unsigned)r.j >= (unsigned)n1) r.j = 0;
    fl = floor(Z);
    r.k = (int)fl;
    r.fz = Z - fl;
    if ((unsigned)r.k >= (unsigned)n2) r.k = 0;
    return r;
}

// CIC axis indices: node origin (j1 = cell) and half-cell origin (j1 = floor(Y - 1/2)).
__device__ __forceinline__ void ax_node(int j, double f, int n, int &j1, int &j2, double &fr) {
    j1 = j;
    j2 = j + 1 == n ? 0 : j + 1;
    fr = f;
}
__device__ __forceinline__ void ax_half(int j, double f, int n, int &j1, int &j2, double &fr) {
    if (f >= 0.5) {
        j1 = j;
        fr = f - 0.5;
    } else {
        j1 = j == 0 ? n - 1 : j - 1;
        fr = f + 0.5;
    }
    j2 = j1 + 1 == n ? 0 : j1 + 1;
}

// Node-origin gather (array_to_particle, src/array_utils.cpp:167-188).
__device__ __forceinline__ double mx_gather(const double *a, const Pos &q, int n1, int n2) {
    int j1, j2, k1, k2;
    double yf, zf;
    ax_node(q.j, q.fy, n1, j1, j2, yf);
    ax_node(q.k, q.fz, n2, k1, k2, zf);
    return gather4u(a, n2, j1, j2, k1, k2, yf, zf);
}
// Gathers on the staggered B grids: B_y has origin (0, dz/2), B_z (dy/2, 0).
__device__ __forceinline__ double mx_gather_hz(const double *a, const Pos &q, int n1, int n2) {
    int j1, j2, k1, k2;
    double yf, zf;
    ax_node(q.j, q.fy, n1, j1, j2, yf);
    ax_half(q.k, q.fz, n2, k1, k2, zf);
    return gather4u(a, n2, j1, j2, k1, k2, yf, zf);
}
__device__ __forceinline__ double mx_gather_hy(const double *a, const Pos &q, int n1, int n2) {
    int j1, j2, k1, k2;
    double yf, zf;
    ax_half(q.j, q.fy, n1, j1, j2, yf);
    ax_node(q.k, q.fz, n2, k1, k2, zf);
    return gather4u(a, n2, j1, j2, k1, k2, yf, zf);
}
// Node value and both staggered derivatives of a node-origin grid at q: the
// gather of mx_gather and the derivative gathers of src/array_utils.cpp:240-288
// (half-node differences (f(j+1) - f(j))/d at j1 = floor(y/d - 1/2), bilinear
// between the two half nodes and the two nodes of the other axis; the
// differences are interpolated first and scaled once by 1/d).  The three
// stencils share eight nodes: the cell's 2x2 corners (c), the row beyond them
// that the y-derivative stencil adds (e_y: row j + 2 when the half-node stencil
// starts at j, else row j - 1) and the column the z-derivative stencil adds
// (e_z), so they take eight loads instead of 4 + 6 + 6.
struct ValGrad {
    double v, dy, dz;
};
// Value and staggered derivatives from the eight nodes (cell corners c, the
// y-stencil's extra row e_y, the z-stencil's extra column e_z).
__device__ __forceinline__ ValGrad vg_math(double c11, double c21, double c12, double c22, double ey1, double ey2,
                                           double ez1, double ez2, double yf, double zf, bool hy, bool hz,
                                           double inv_d1, double inv_d2) {
    ValGrad o;
    o.v = bilin(c11, c21, c12, c22, yf, zf);
    // Derivatives: along each line of the other axis the half-node stencil
    // interpolates between the middle difference (at the cell's half node) and
    // the outer one (at the half node beyond the cell on the particle's side),
    // with weight t = |f - 1/2| on the outer one.  That is the reference's
    // two-half-node interpolation with its node choice folded into one operand
    // select; it agrees with the stencil as written to rounding (7e-16 of the
    // stencil's largest node over 2e5 random cases on the CPU).
    {  // y: columns k1, k2
        const double t = fabs(yf - 0.5);
        const double d1 = c21 - c11, d2 = c22 - c12;
        const double s1 = hy ? ey1 - c21 : c11 - ey1, s2 = hy ? ey2 - c22 : c12 - ey2;
        const double g1 = d1 + (s1 - d1) * t, g2 = d2 + (s2 - d2) * t;
        o.dy = (g1 * (1 - zf) + g2 * zf) * inv_d1;
    }
    {  // z: rows j1, j2
        const double t = fabs(zf - 0.5);
        const double d1 = c12 - c11, d2 = c22 - c21;
        const double s1 = hz ? ez1 - c12 : c11 - ez1, s2 = hz ? ez2 - c22 : c21 - ez2;
        const double g1 = d1 + (s1 - d1) * t, g2 = d2 + (s2 - d2) * t;
        o.dz = (g1 * (1 - yf) + g2 * yf) * inv_d2;
    }
    return o;
}
__device__ __forceinline__ ValGrad mx_gather_vg(const double *a, const Pos &q, int n1, int n2, double inv_d1,
                                                double inv_d2) {
    const int j1 = q.j, j2 = q.j + 1 == n1 ? 0 : q.j + 1;
    const int k1 = q.k, k2 = q.k + 1 == n2 ? 0 : q.k + 1;
    const bool hy = q.fy >= 0.5, hz = q.fz >= 0.5;
    const int jx = hy ? (j2 + 1 == n1 ? 0 : j2 + 1) : (j1 == 0 ? n1 - 1 : j1 - 1);
    const int kx = hz ? (k2 + 1 == n2 ? 0 : k2 + 1) : (k1 == 0 ? n2 - 1 : k1 - 1);
    const unsigned r1 = __umul24((unsigned)j1, (unsigned)n2), r2 = __umul24((unsigned)j2, (unsigned)n2),
                   rx = __umul24((unsigned)jx, (unsigned)n2);
    return vg_math(ld_at(a, r1 + k1), ld_at(a, r2 + k1), ld_at(a, r1 + k2), ld_at(a, r2 + k2), ld_at(a, rx + k1),
                   ld_at(a, rx + k2), ld_at(a, r1 + kx), ld_at(a, r2 + kx), q.fy, q.fz, hy, hz, inv_d1, inv_d2);
}

// ------------------------------------------------------ LDS deposit window
// Each block deposits into an LDS window covering the nodes its particles
// touch, then flushes the window with one global atomic per non-zero node.
// The window is found from the block's own particles (periodic-aware, relative
// to the first particle's cell).  If it exceeds CAP nodes the block falls
// back to direct global atomics.
struct Win {
    int j0, k0, wj, wk;
    bool lds;
};

__device__ __forceinline__ int rel_idx(int j, int ref, int n) {
    int d = j - ref;
    if (d < -(n / 2)) d += n;
    else if (d >= n - n / 2) d -= n;
    return d;
}

// Sum over the block (blockDim.x a multiple of 64, <= 1024); result in every thread.
__device__ inline double block_sum(double v) {
    __shared__ double part[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) part[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += part[i];
    return s;
}

// Min over the wavefront, the same value in every lane: four DPP steps inside
// each 16-lane row (lane ^ 1, lane ^ 2, half-row mirror, row mirror) leave the
// row's min in all of its lanes, then the four rows are read out.  Every lane
// of the wavefront must call it.
// (The compiler folds these moves into the min / max that uses them.)
template <int CTRL>
__device__ __forceinline__ int dpp_row(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int wave_min(int v) {
    v = min(v, dpp_row<0xB1>(v));   // quad_perm(1, 0, 3, 2)
    v = min(v, dpp_row<0x4E>(v));   // quad_perm(2, 3, 0, 1)
    v = min(v, dpp_row<0x141>(v));  // row_half_mirror
    v = min(v, dpp_row<0x140>(v));  // row_mirror
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}
__device__ __forceinline__ int wave_max(int v) {
    v = max(v, dpp_row<0xB1>(v));
    v = max(v, dpp_row<0x4E>(v));
    v = max(v, dpp_row<0x141>(v));
    v = max(v, dpp_row<0x140>(v));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// rmin/rmax/cmin/cmax: this thread's relative node range (INT_MAX / INT_MIN
// when the thread deposits nothing).  Relative indices come from rel_idx, so
// they lie in [-n/2, n - n/2].
// `sh` is 4*(blockDim/64)+2 ints of LDS.  The window is the same in every lane
// (kept in scalar registers).
template <int CAP>
__device__ inline Win block_window(int rmin, int rmax, int cmin, int cmax, int jref, int kref, int n1, int n2,
                                   int *sh) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    rmin = wave_min(rmin);
    rmax = wave_max(rmax);
    cmin = wave_min(cmin);
    cmax = wave_max(cmax);
    if (lane == 0) {
        sh[4 * wv + 0] = rmin;
        sh[4 * wv + 1] = rmax;
        sh[4 * wv + 2] = cmin;
        sh[4 * wv + 3] = cmax;
    }
    __syncthreads();
    int a = sh[0], b = sh[1], c = sh[2], d = sh[3];
    for (int i = 1; i < nw; ++i) {
        a = min(a, sh[4 * i + 0]);
        b = max(b, sh[4 * i + 1]);
        c = min(c, sh[4 * i + 2]);
        d = max(d, sh[4 * i + 3]);
    }
    __syncthreads();
    a = __builtin_amdgcn_readfirstlane(a);
    b = __builtin_amdgcn_readfirstlane(b);
    c = __builtin_amdgcn_readfirstlane(c);
    d = __builtin_amdgcn_readfirstlane(d);
    Win w;
    if (a > b) {  // no deposits in this block
        w.j0 = w.k0 = 0;
        w.wj = w.wk = 0;
        w.lds = true;
        return w;
    }
    w.wj = b - a + 1;
    w.wk = d - c + 1;
    // jref in [0, n1) and a in [-n1/2, n1 - n1/2]: one correction wraps j0
    w.j0 = __builtin_amdgcn_readfirstlane(jref) + a;
    if (w.j0 < 0) w.j0 += n1;
    else if (w.j0 >= n1) w.j0 -= n1;
    w.k0 = __builtin_amdgcn_readfirstlane(kref) + c;
    if (w.k0 < 0) w.k0 += n2;
    else if (w.k0 >= n2) w.k0 -= n2;
    w.lds = (w.wj <= n1 && w.wk <= n2 && w.wj * w.wk <= CAP);
    return w;
}

// Count a block that deposits but whose window exceeds its capacity (it falls back
// to global atomics): one atomic per such block, none otherwise.
__device__ __forceinline__ void win_note_overflow(const Win &w, Ctl *ctl) {
    if (!w.lds && w.wj > 0 && threadIdx.x == 0) atomicAdd(&ctl->ovf, 1u);
}

// Window slot of node (j, k) for a CIC stencil whose other nodes are (j + 1, k),
// (j, k + 1), (j + 1, k + 1) (periodic): those sit at + wk, + 1 and + wk + 1,
// since the stencil stays inside the window.  -1 when it does not (the
// deposit then goes to global memory).
__device__ __forceinline__ int win_base(const Win &w, int j, int k, int n1, int n2) {
    int dj = j - w.j0;
    if (dj < 0) dj += n1;
    int dk = k - w.k0;
    if (dk < 0) dk += n2;
    return (dj + 1 < w.wj && dk + 1 < w.wk) ? dj * w.wk + dk : -1;
}

// ------------------------------------------------ quad folding of deposits
// Particles are stored cell by cell, so the lanes of a quad (4 consecutive
// lanes) often deposit onto the same four nodes.  Folding their weights into
// one lane before the atomics cuts the LDS atomics and their same-address
// conflicts by up to 4x.  Lane exchange by DPP quad permutations.
// A quad permutation reads a valid lane for every lane, so no "old" value is
// needed: mov_dpp is one instruction, without a zeroed destination first.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffLL));
    const int hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// key >= 0 identifies the stencil (equal keys = same four target nodes); on
// return the lowest lane of every group of quad lanes with equal keys holds
// the summed weights and keeps `act`, the others have act = false.  Every lane
// of the wavefront must call it (no divergent branch around the call).
template <int CTRL, int BIT, int NV>
__device__ __forceinline__ void quad_fold_step(int key, bool &act, double (&a)[NV]) {
    const int lane = threadIdx.x & 63;
    const int k = act ? key : -1 - lane;
    const int pk = dpp_i<CTRL>(k);
    double pa[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) pa[i] = dpp_d<CTRL>(a[i]);
    if (act && pk == k) {
        if (lane & BIT) {
            act = false;
        } else {
#pragma unroll
            for (int i = 0; i < NV; ++i) a[i] += pa[i];
        }
    }
}
template <int NV>
__device__ __forceinline__ void quad_fold(int key, bool &act, double (&a)[NV]) {
    quad_fold_step<0xB1, 1, NV>(key, act, a);  // quad_perm(1, 0, 3, 2): lane ^ 1
    quad_fold_step<0x4E, 2, NV>(key, act, a);  // quad_perm(2, 3, 0, 1): lane ^ 2
}

// CIC weights of value v at fractions (yf, zf): nodes (1,1), (2,1), (1,2), (2,2).
__device__ __forceinline__ void cic_w(double v, double yf, double zf, double *a) {
    a[0] = v * (1 - yf) * (1 - zf);
    a[1] = v * yf * (1 - zf);
    a[2] = v * (1 - yf) * zf;
    a[3] = v * yf * zf;
}

// LDS-typed pointer: atomics through it are ds_add_f64.  With a generic
// pointer the compiler may merge the LDS and global branches of dep4w into one
// flat atomic, which is much slower on LDS.
typedef __attribute__((address_space(3))) double lds_f64;

// One CIC deposit of weights a11..a22 at (j1..j2, k1..k2) into a grid (LDS
// window or global); j2 = j1 + 1 and k2 = k1 + 1 up to the periodic wrap.
__device__ __forceinline__ void dep4w(double *lds, const Win &w, double *glob, int n1, int n2, int j1, int j2,
                                      int k1, int k2, double a11, double a21, double a12, double a22) {
    if (w.lds) {
        const int s = win_base(w, j1, k1, n1, n2);
        if (s >= 0) {
            lds_f64 *l = (lds_f64 *)lds + s;
            __hip_atomic_fetch_add(l, a11, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(l + w.wk, a21, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(l + 1, a12, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(l + w.wk + 1, a22, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return;
        }
    }
    unsafeAtomicAdd(at(glob, nidx(j1, k1, n2)), a11);
    unsafeAtomicAdd(at(glob, nidx(j2, k1, n2)), a21);
    unsafeAtomicAdd(at(glob, nidx(j1, k2, n2)), a12);
    unsafeAtomicAdd(at(glob, nidx(j2, k2, n2)), a22);
}

// G grids sharing one stencil (weights a[4 g .. 4 g + 3] for grid g, LDS
// windows G apart by CAP): one slot computation and one branch for all.
template <int G, int CAP>
__device__ __forceinline__ void dep4w_multi(double *lds, const Win &w, double *const *glob, int n1, int n2, int j1,
                                            int j2, int k1, int k2, const double *a) {
    if (w.lds) {
        const int s = win_base(w, j1, k1, n1, n2);
        if (s >= 0) {
            lds_f64 *l = (lds_f64 *)lds + s;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                __hip_atomic_fetch_add(l + g * CAP, a[4 * g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(l + g * CAP + w.wk, a[4 * g + 1], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(l + g * CAP + 1, a[4 * g + 2], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_add(l + g * CAP + w.wk + 1, a[4 * g + 3], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            return;
        }
    }
    const unsigned t11 = nidx(j1, k1, n2), t21 = nidx(j2, k1, n2), t12 = nidx(j1, k2, n2), t22 = nidx(j2, k2, n2);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        unsafeAtomicAdd(at(glob[g], t11), a[4 * g]);
        unsafeAtomicAdd(at(glob[g], t21), a[4 * g + 1]);
        unsafeAtomicAdd(at(glob[g], t12), a[4 * g + 2]);
        unsafeAtomicAdd(at(glob[g], t22), a[4 * g + 3]);
    }
}

template <int CAP>
__device__ inline void win_zero(double *lds, int ngrid, const Win &w) {
    if (!w.lds) return;
    const int m = w.wj * w.wk;
    for (int g = 0; g < ngrid; ++g)
        for (int i = threadIdx.x; i < m; i += blockDim.x) lds[g * CAP + i] = 0.0;
}

// Row of window slot i = dj * wk + dk: (i + 1/2) / wk sits at least 1/(2 wk)
// from an integer, and the single-precision product errs by less than
// 2^-22 CAP < 1/(2 CAP) for i < CAP, wk <= CAP (CAP < 1448).
template <int CAP>
__device__ inline void win_flush(const double *lds, double *const *glob, int ngrid, const Win &w, int n1, int n2) {
    static_assert(CAP < 1448, "window slot rows come from a single-precision reciprocal");
    if (!w.lds) return;
    const int m = w.wj * w.wk;
    const float inv_wk = 1.0f / (float)w.wk;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const int dj = (int)(((float)i + 0.5f) * inv_wk), dk = i - dj * w.wk;
        int j = w.j0 + dj;
        if (j >= n1) j -= n1;
        int k = w.k0 + dk;
        if (k >= n2) k -= n2;
        const unsigned t = nidx(j, k, n2);
        for (int g = 0; g < ngrid; ++g) {
            const double v = lds[g * CAP + i];
            if (v != 0.0) unsafeAtomicAdd(at(glob[g], t), v);
        }
    }
}

// Relative node range [r, r+1] x [c, c+1] of a deposit at (j1, k1).
__device__ __forceinline__ void range_add(int j1, int k1, int jref, int kref, int n1, int n2, int &rmin, int &rmax,
                                          int &cmin, int &cmax) {
    const int r = rel_idx(j1, jref, n1), c = rel_idx(k1, kref, n2);
    rmin = min(rmin, r);
    rmax = max(rmax, r + 1);
    cmin = min(cmin, c);
    cmax = max(cmax, c + 1);
}

// ------------------------------------------------------ LDS gather window
// The gathers of a block's particles read the grid nodes around the block's
// cells.  Issued per particle they are scattered 8-byte loads, and those keep
// the CU's vector-memory address and data units (TA, TD) busy 80-90 % of a push
// launch at 1024^2 (counters), which bounds the kernel.  Staged instead: the
// block copies the window of nodes its stencils can touch into LDS with
// row-contiguous loads, and every gather reads LDS.  Window rows and columns are
// relative to its origin and cover each stencil, so no periodic wrap is needed
// inside; the arithmetic is that of the global gathers, so both give the same
// bits.  A block whose window exceeds the capacity gathers from global memory.
// Nodes per staged grid (rows x columns of the window); 384 covers a 256-particle
// block of a sorted 4-per-cell lattice (about 7 x 19 nodes) with room for motion.
constexpr int kGwCap = 384;
static_assert(kGwCap < 1448, "window slot rows come from a single-precision reciprocal");
template <int NG>
__device__ inline void gwin_load(double *gw, int gcap, const double *const (&src)[NG], const Win &w, int n1,
                                 int n2) {
    const int m = w.wj * w.wk;
    const float inv_wk = 1.0f / (float)w.wk;  // exact rows for m < 1448 (as in win_flush)
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const int dj = (int)(((float)i + 0.5f) * inv_wk), dk = i - dj * w.wk;
        int j = w.j0 + dj;
        if (j >= n1) j -= n1;
        int k = w.k0 + dk;
        if (k >= n2) k -= n2;
        const unsigned t = nidx(j, k, n2);
#pragma unroll
        for (int g = 0; g < NG; ++g) gw[g * gcap + i] = ld_at(src[g], t);
    }
}
// Window-relative row and column of cell (j, k).
__device__ __forceinline__ void gwin_rc(const Win &w, int j, int k, int n1, int n2, int &r, int &c) {
    r = j - w.j0;
    if (r < 0) r += n1;
    c = k - w.k0;
    if (c < 0) c += n2;
}
// Node-origin gather at window cell (r, c).
__device__ __forceinline__ double lg_node(const double *l, int wk, int r, int c, double yf, double zf) {
    const double *l1 = l + r * wk + c;
    return bilin(l1[0], l1[wk], l1[1], l1[wk + 1], yf, zf);
}
// B_y on (0, dz/2): node rows, half-node columns (ax_node x ax_half).
__device__ __forceinline__ double lg_hz(const double *l, int wk, int r, int c, double fy, double fz) {
    const bool h = fz >= 0.5;
    return lg_node(l, wk, r, h ? c : c - 1, fy, h ? fz - 0.5 : fz + 0.5);
}
// B_z on (dy/2, 0): half-node rows, node columns (ax_half x ax_node).
__device__ __forceinline__ double lg_hy(const double *l, int wk, int r, int c, double fy, double fz) {
    const bool h = fy >= 0.5;
    return lg_node(l, wk, h ? r : r - 1, c, h ? fy - 0.5 : fy + 0.5, fz);
}
__device__ __forceinline__ ValGrad lg_vg(const double *l, int wk, int r, int c, double fy, double fz, double inv_d1,
                                         double inv_d2) {
    const bool hy = fy >= 0.5, hz = fz >= 0.5;
    const double *l1 = l + r * wk, *l2 = l1 + wk, *lx = l + (hy ? r + 2 : r - 1) * wk;
    const int cx = hz ? c + 2 : c - 1;
    return vg_math(l1[c], l2[c], l1[c + 1], l2[c + 1], lx[c], lx[c + 1], l1[cx], l2[cx], fy, fz, hy, hz, inv_d1,
                   inv_d2);
}

}  // namespace qp
