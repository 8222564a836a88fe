// qp_bunch.hip -- the driver bunch on the device: streamed per-slice sources
// and field gathers for the wake loops, the explicit push and re-bucketing
// (src/System_3d.cpp:843-946, CIC of src/array_utils.cpp:290-374, field
// array origins of src/System_3d.cpp:179-183).
//
// Arithmetic follows the reference's operation order and this file is built
// without FMA contraction, like the plain FP64 path, so that deposits (up to
// the order of the atomic sums), gathers and the push agree with the oracle at
// rounding level.  Index wrap: the reference computes j2 = (j1+1) % n before
// wrapping a negative j1 once (src/array_utils.cpp:303-314); both indices are
// wrapped fully here, which is the same wherever the reference stays in
// bounds.  Not restated on the device: QED photon emission (:892-921; random,
// from the unseeded rand()).
#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include "qp_bunch.h"

namespace qp {

namespace {

constexpr int kBT = 256;
constexpr int kSliceBlocks = 128;  // grid of the per-slice launches (bucket sizes vary per slice)

inline int bblk(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + kBT - 1) / kBT, 65535 * 4)); }

__device__ __forceinline__ int wrap_idx(int j, int n) {
    j %= n;
    return j < 0 ? j + n : j;
}

__device__ __forceinline__ int slice_index(const Ctl *ctl, int slot, int slice) {
    return slot >= 0 ? ctl->slice_of[slot] : slice;
}

// Transverse CIC indices and fractions on the nodes (origin 0).
struct Tw {
    int j1, j2, k1, k2;
    double yf, zf;
};
__device__ __forceinline__ Tw tw_of(double y, double z, const BunchDev &B, double oy = 0.0, double oz = 0.0) {
    Tw t;
    const int j = (int)floor((y - oy) / B.dy);
    t.yf = (y - oy) / B.dy - j;
    t.j2 = wrap_idx(j + 1, B.n1);
    t.j1 = wrap_idx(j, B.n1);
    const int k = (int)floor((z - oz) / B.dz);
    t.zf = (z - oz) / B.dz - k;
    t.k2 = wrap_idx(k + 1, B.n2);
    t.k1 = wrap_idx(k, B.n2);
    return t;
}

// Buckets are half x-cells: h = floor(2x/dx), bucket h + 1 for h in
// [-1, 2nx - 3] (every particle some array of the reference deposits to or
// gathers from: the sources and most fields sit on the nodes, e_x on nodes
// shifted by -dx/2), bucket 2nx - 1 for the rest.  floor(h/2) is the node
// x-cell floor(x/dx) exactly (2x/dx is exactly twice x/dx).
__device__ __forceinline__ int bucket_of(double x, const BunchDev &B) {
    const double h = floor(2.0 * x / B.dx);
    return (h >= -1.0 && h <= (double)(2 * B.nx - 3)) ? (int)h + 1 : 2 * B.nx - 1;
}

// Particle range of the buckets [b0, b1] (clipped to the valid buckets).
__device__ __forceinline__ void bucket_range(const BunchDev &B, int b0, int b1, int &p0, int &np) {
    b0 = b0 < 0 ? 0 : b0;
    b1 = b1 > 2 * B.nx - 2 ? 2 * B.nx - 2 : b1;
    if (b1 < b0) {
        p0 = np = 0;
        return;
    }
    p0 = B.start[b0];
    np = B.start[b1 + 1] - p0;
}

// Slot of this lane in counter ctr[b], one atomic per distinct b in the
// wavefront (the bunch is sorted by bucket, so a wavefront mostly holds one or
// two buckets and per-particle atomics would all hit the same counter).  Every
// lane of the wavefront calls it; inactive lanes take no slot.
__device__ __forceinline__ unsigned wave_slot(unsigned *ctr, int b, bool active) {
    const int lane = threadIdx.x & 63;
    bool done = !active;
    unsigned res = 0;
    while (true) {
        const unsigned long long pending = __ballot(!done);
        if (!pending) break;
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lk = __shfl(b, leader);
        const bool mine = !done && b == lk;
        const unsigned long long m = __ballot(mine);
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&ctr[lk], (unsigned)__popcll(m));
        base = (unsigned)__shfl((int)base, leader);
        if (mine) {
            res = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
            done = true;
        }
    }
    return res;
}

__global__ void k_bunch_zero(double *planes, size_t n, const Ctl *ctl) {
    if (ctl && ctl->stop_acc) return;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        planes[t] = 0.0;
}

// rho_b(i), jx_b(i) from the particles whose x-cell i1 = floor(x/dx) is i
// (weight 1 - x_frac) or i - 1 (weight x_frac), inside the reference's x cut
// 0 <= i1, i1 + 1 < nx (src/array_utils.cpp:290-338): half-buckets 2i-1..2i+2.
__global__ __launch_bounds__(kBT) void k_bunch_src(BunchDev B, double *rho, double *jx, const Ctl *ctl, int slot,
                                                   int slice) {
    if (ctl && ctl->stop_acc) return;
    const int i = slice_index(ctl, slot, slice);
    if (i < 0 || i >= B.nx) return;
    int p0, np;
    bucket_range(B, 2 * i - 1, 2 * i + 2, p0, np);
    const int n2 = B.n2;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x) {
        const int p = p0 + t;
        const double x = B.x[p];
        const int i1 = (int)floor(x / B.dx);
        if (i1 < 0 || i1 + 1 >= B.nx || (i1 != i && i1 + 1 != i)) continue;
        const double x_frac = x / B.dx - i1;
        const Tw w = tw_of(B.y[p], B.z[p], B);
        const double v = B.q[p];
        const double vj = v * B.px[p] / B.gamma[p];
        const double wx = i1 == i ? (1 - x_frac) : x_frac;
        const size_t c11 = (size_t)w.j1 * n2 + w.k1, c21 = (size_t)w.j2 * n2 + w.k1;
        const size_t c12 = (size_t)w.j1 * n2 + w.k2, c22 = (size_t)w.j2 * n2 + w.k2;
        unsafeAtomicAdd(&rho[c11], v * wx * (1 - w.yf) * (1 - w.zf));
        unsafeAtomicAdd(&rho[c21], v * wx * w.yf * (1 - w.zf));
        unsafeAtomicAdd(&rho[c12], v * wx * (1 - w.yf) * w.zf);
        unsafeAtomicAdd(&rho[c22], v * wx * w.yf * w.zf);
        unsafeAtomicAdd(&jx[c11], vj * wx * (1 - w.yf) * (1 - w.zf));
        unsafeAtomicAdd(&jx[c21], vj * wx * w.yf * (1 - w.zf));
        unsafeAtomicAdd(&jx[c12], vj * wx * (1 - w.yf) * w.zf);
        unsafeAtomicAdd(&jx[c22], vj * wx * w.yf * w.zf);
    }
}

// The two halves of array_to_particle (src/array_utils.cpp:340-374): the
// four terms of the x-cell's lower slice i1 (weight 1 - x_frac) and of its
// upper slice i2 = i1 + 1 (weight x_frac), from one 2D plane each.
__device__ __forceinline__ double half_lo(const double *a, const Tw &w, int n2, double x_frac) {
    return a[(size_t)w.j1 * n2 + w.k1] * (1 - x_frac) * (1 - w.yf) * (1 - w.zf) +
           a[(size_t)w.j2 * n2 + w.k1] * (1 - x_frac) * w.yf * (1 - w.zf) +
           a[(size_t)w.j1 * n2 + w.k2] * (1 - x_frac) * (1 - w.yf) * w.zf +
           a[(size_t)w.j2 * n2 + w.k2] * (1 - x_frac) * w.yf * w.zf;
}
__device__ __forceinline__ double half_hi(const double *b, const Tw &w, int n2, double x_frac) {
    return b[(size_t)w.j1 * n2 + w.k1] * x_frac * (1 - w.yf) * (1 - w.zf) +
           b[(size_t)w.j2 * n2 + w.k1] * x_frac * w.yf * (1 - w.zf) +
           b[(size_t)w.j1 * n2 + w.k2] * x_frac * (1 - w.yf) * w.zf +
           b[(size_t)w.j2 * n2 + w.k2] * x_frac * w.yf * w.zf;
}

struct Planes5 {
    const double *f[5];  // ex, ey, ez, by, bz of the slice
};

// Once slice i's fields are final, every particle for which slice i is a face
// of its x-cell in some field array takes that half of the field's trilinear
// gather.  Each field is read with its array's origin (src/System_3d.cpp:179-183:
// e_x (-dx/2, 0, 0), e_y and B_z (0, dy/2, 0), e_z and B_y (0, 0, dz/2)), so
// its x-cell i1 = floor((x - o_x)/dx) is the node cell or the one above.  The
// sweep reaches the upper face i1 + 1 first: that half is stored in the field
// array, and the lower half at slice i1 is added to it.  chi
// (src/System_3d.cpp:843-860) is recomputed at every visit; the last visit
// completes all five fields.  Half-buckets 2i-3..2i+2 hold every such particle
// (with a margin for the rounding of (x + dx/2)/dx).
__global__ __launch_bounds__(kBT) void k_bunch_gather(BunchDev B, Planes5 cur, const Ctl *ctl, int slot,
                                                      int slice) {
    if (ctl && ctl->stop_acc) return;
    const int i = slice_index(ctl, slot, slice);
    if (i < 0 || i >= B.nx) return;
    int p0, np;
    bucket_range(B, 2 * i - 3, 2 * i + 2, p0, np);
    const int n2 = B.n2;
    const double hx = 0.5 * B.dx, hy = 0.5 * B.dy, hz = 0.5 * B.dz;
    const double ox[5] = {-hx, 0.0, 0.0, 0.0, 0.0}, oy[5] = {0.0, hy, 0.0, 0.0, hy}, oz[5] = {0.0, 0.0, hz, hz, 0.0};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < np; t += gridDim.x * blockDim.x) {
        const int p = p0 + t;
        const double x = B.x[p], y = B.y[p], z = B.z[p];
        double *out[5] = {B.ex, B.ey, B.ez, B.by, B.bz};
        bool touched = false;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const int i1 = (int)floor((x - ox[f]) / B.dx);
            if (i1 < 0 || i1 + 1 >= B.nx || (i1 != i && i1 + 1 != i)) continue;
            const double x_frac = (x - ox[f]) / B.dx - i1;
            const Tw w = tw_of(y, z, B, oy[f], oz[f]);
            if (i1 + 1 == i)
                out[f][p] = half_hi(cur.f[f], w, n2, x_frac);
            else
                out[f][p] = half_lo(cur.f[f], w, n2, x_frac) + out[f][p];
            touched = true;
        }
        if (!touched) continue;
        const double ex = B.ex[p], ey = B.ey[p], ez = B.ez[p], by = B.by[p], bz = B.bz[p];
        const double px = B.px[p], py = B.py[p], pz = B.pz[p], g = B.gamma[p];
        const double fx = g * ex + py * bz - pz * by;
        const double fy = g * ey - px * bz;
        const double fz = g * ez + px * by;
        const double fl = px * ex + py * ey + pz * ez;
        B.chi[p] = sqrt(fx * fx + fy * fy + fz * fz - fl * fl) / B.fs;
    }
}

__global__ void k_bunch_clear(BunchDev B) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= B.np) return;
    B.ex[p] = B.ey[p] = B.ez[p] = B.by[p] = B.bz[p] = B.chi[p] = 0.0;
}

// advance_bunch_particles without QED emission (src/System_3d.cpp:862-890):
// reads S, writes the advanced particle to D (same slot) with its bucket.
__device__ __forceinline__ int push_one(const BunchDev &S, const BunchDev &D, double dt, int p) {
    double x = S.x[p], y = S.y[p], z = S.z[p], px = S.px[p], py = S.py[p], pz = S.pz[p], g = S.gamma[p];
    const double ex = S.ex[p], ey = S.ey[p], ez = S.ez[p], by = S.by[p], bz = S.bz[p];
    if (S.photon[p]) {
        x += dt * (px / g - 1);
        y += dt * py / g;
        z += dt * pz / g;
        g = sqrt(px * px + py * py + pz * pz);
    } else {
        const double cmr = S.cmr[p];
        const double fx = cmr * (ex + (py * bz - pz * by) / g);
        const double fy = cmr * (ey - px * bz / g);
        const double fz = cmr * (ez + px * by / g);
        x += dt * (px / g - 1);
        y += dt * py / g;
        z += dt * pz / g;
        px += dt * fx;
        py += dt * fy;
        pz += dt * fz;
        g = sqrt(1 + px * px + py * py + pz * pz);
    }
    D.x[p] = x;
    D.y[p] = y;
    D.z[p] = z;
    D.px[p] = px;
    D.py[p] = py;
    D.pz[p] = pz;
    D.gamma[p] = g;
    D.q[p] = S.q[p];
    D.cmr[p] = S.cmr[p];
    D.photon[p] = S.photon[p];
    D.id[p] = S.id[p];
    D.ex[p] = ex;
    D.ey[p] = ey;
    D.ez[p] = ez;
    D.by[p] = by;
    D.bz[p] = bz;
    D.chi[p] = S.chi[p];
    return bucket_of(x, D);
}

__global__ void k_bunch_push(BunchDev S, BunchDev D, double dt, int *key, unsigned *count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < S.np;
    int b = 0;
    if (valid) b = push_one(S, D, dt, p);
    (void)wave_slot(count, b, valid);
    if (valid) key[p] = b;
}

__global__ void k_bunch_key(BunchDev S, int *key, unsigned *count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < S.np;
    const int b = valid ? bucket_of(S.x[p], S) : 0;
    (void)wave_slot(count, b, valid);
    if (valid) key[p] = b;
}

// Exclusive scan of nb bucket counts into start[0..nb] (one block of 1024).
__global__ void k_bunch_scan(const unsigned *count, int *start, unsigned *cursor, int nb) {
    __shared__ unsigned part[1024];
    const int per = (nb + blockDim.x - 1) / blockDim.x;
    const int b0 = threadIdx.x * per;
    unsigned s = 0;
    for (int i = b0; i < b0 + per && i < nb; ++i) s += count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            const unsigned v = part[i];
            part[i] = run;
            run += v;
        }
        start[nb] = (int)run;
    }
    __syncthreads();
    unsigned run = part[threadIdx.x];
    for (int i = b0; i < b0 + per && i < nb; ++i) {
        start[i] = (int)run;
        cursor[i] = 0;
        run += count[i];
    }
}

__global__ void k_bunch_scatter(BunchDev S, BunchDev D, const int *key, unsigned *cursor) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < S.np;
    const int b = valid ? key[p] : 0;
    const unsigned slot = wave_slot(cursor, b, valid);
    if (!valid) return;
    const unsigned d = (unsigned)D.start[b] + slot;
    if (d >= (unsigned)S.np) return;
    D.x[d] = S.x[p];
    D.y[d] = S.y[p];
    D.z[d] = S.z[p];
    D.px[d] = S.px[p];
    D.py[d] = S.py[p];
    D.pz[d] = S.pz[p];
    D.gamma[d] = S.gamma[p];
    D.q[d] = S.q[p];
    D.cmr[d] = S.cmr[p];
    D.photon[d] = S.photon[p];
    D.id[d] = S.id[p];
    D.ex[d] = S.ex[p];
    D.ey[d] = S.ey[p];
    D.ez[d] = S.ez[p];
    D.by[d] = S.by[p];
    D.bz[d] = S.bz[p];
    D.chi[d] = S.chi[p];
}

// Full 3D deposit (deposit_bunch_densities, src/System_3d.cpp:926-946) into
// nx*ny*nz arrays: the reference's 3D output of rho_bunch and jx_bunch.
__global__ void k_bunch_dep3d(BunchDev B, double *rho3, double *jx3) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= B.np) return;
    const double x = B.x[p];
    const int i1 = (int)floor(x / B.dx), i2 = i1 + 1;
    if (i1 < 0 || i2 >= B.nx) return;
    const double x_frac = x / B.dx - i1;
    const Tw w = tw_of(B.y[p], B.z[p], B);
    const int n2 = B.n2;
    const size_t nn = (size_t)B.n1 * n2;
    const size_t c11 = (size_t)w.j1 * n2 + w.k1, c21 = (size_t)w.j2 * n2 + w.k1;
    const size_t c12 = (size_t)w.j1 * n2 + w.k2, c22 = (size_t)w.j2 * n2 + w.k2;
    const double vals[2] = {B.q[p], B.q[p] * B.px[p] / B.gamma[p]};
    double *arr[2] = {rho3, jx3};
    for (int a = 0; a < 2; ++a) {
        const double v = vals[a];
        double *lo = arr[a] + (size_t)i1 * nn, *hi = arr[a] + (size_t)i2 * nn;
        unsafeAtomicAdd(&lo[c11], v * (1 - x_frac) * (1 - w.yf) * (1 - w.zf));
        unsafeAtomicAdd(&lo[c21], v * (1 - x_frac) * w.yf * (1 - w.zf));
        unsafeAtomicAdd(&lo[c12], v * (1 - x_frac) * (1 - w.yf) * w.zf);
        unsafeAtomicAdd(&lo[c22], v * (1 - x_frac) * w.yf * w.zf);
        unsafeAtomicAdd(&hi[c11], v * x_frac * (1 - w.yf) * (1 - w.zf));
        unsafeAtomicAdd(&hi[c21], v * x_frac * w.yf * (1 - w.zf));
        unsafeAtomicAdd(&hi[c12], v * x_frac * (1 - w.yf) * w.zf);
        unsafeAtomicAdd(&hi[c22], v * x_frac * w.yf * w.zf);
    }
}

#define BLAUNCH(kern, grid, block, ...)                                       \
    do {                                                                      \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__); \
        QP_HIP_CHECK(hipGetLastError());                                      \
    } while (0)

void free_set(BunchDev &d) {
    void *arrs[] = {d.x, d.y, d.z, d.px, d.py, d.pz, d.gamma, d.q, d.cmr, d.photon, d.id,
                    d.ex, d.ey, d.ez, d.by, d.bz, d.chi};
    for (void *a : arrs)
        if (a) (void)hipFree(a);
    const BunchDev keep = d;
    d = BunchDev{};
    d.start = keep.start;
    d.nx = keep.nx;
    d.n1 = keep.n1;
    d.n2 = keep.n2;
    d.dx = keep.dx;
    d.dy = keep.dy;
    d.dz = keep.dz;
    d.fs = keep.fs;
}

int alloc_set(BunchDev &d, size_t n) {
    double **dbl[] = {&d.x, &d.y, &d.z, &d.px, &d.py, &d.pz, &d.gamma, &d.q, &d.cmr,
                      &d.ex, &d.ey, &d.ez, &d.by, &d.bz, &d.chi};
    for (double **a : dbl) QP_HIP_CHECK(hipMalloc(a, sizeof(double) * n));
    QP_HIP_CHECK(hipMalloc(&d.photon, sizeof(int) * n));
    QP_HIP_CHECK(hipMalloc(&d.id, sizeof(int) * n));
    return QP_OK;
}

// Sort set S (np particles) into b->cur by bucket; key and count already
// filled by the caller's kernel on S.
int bucket_sort(Bunch *b, BunchDev &S, hipStream_t s) {
    BLAUNCH(k_bunch_scan, 1, 1024, (const unsigned *)b->count, b->start, (unsigned *)b->cursor, 2 * b->nx);
    BLAUNCH(k_bunch_scatter, bblk(b->np), kBT, S, b->cur, (const int *)b->key, (unsigned *)b->cursor);
    return QP_OK;
}

}  // namespace

// ------------------------------------------------------------ wake hooks
int bunch_clear_fields(Bunch *b, hipStream_t s) {
    if (b->np) BLAUNCH(k_bunch_clear, bblk(b->np), kBT, b->cur);
    return QP_OK;
}

int bunch_slice_sources(Bunch *b, const Ctl *ctl, int slot, int slice, hipStream_t s) {
    BLAUNCH(k_bunch_zero, bblk(2 * b->nn()), kBT, b->planes, 2 * b->nn(), ctl);
    BLAUNCH(k_bunch_src, kSliceBlocks, kBT, b->cur, b->rho_plane(), b->jx_plane(), ctl, slot, slice);
    return QP_OK;
}

int bunch_slice_gather(Bunch *b, const double *ex, const double *ey, const double *ez, const double *by,
                       const double *bz, const Ctl *ctl, int slot, int slice, hipStream_t s) {
    Planes5 cur{{ex, ey, ez, by, bz}};
    BLAUNCH(k_bunch_gather, kSliceBlocks, kBT, b->cur, cur, ctl, slot, slice);
    return QP_OK;
}

}  // namespace qp

using namespace qp;

#define BARG_CHECK(cond, msg)        \
    do {                             \
        if (!(cond)) {               \
            set_error(msg);          \
            return QP_ERR_ARG;       \
        }                            \
    } while (0)

namespace {

// Host copy of a device array in upload order (the device keeps the bunch sorted).
template <class T>
int unsort(const Bunch *b, const T *dev, T *host, const std::vector<int> &id) {
    if (!host) return QP_OK;
    std::vector<T> tmp(b->np);
    QP_HIP_CHECK(hipMemcpy(tmp.data(), dev, sizeof(T) * b->np, hipMemcpyDeviceToHost));
    for (size_t p = 0; p < b->np; ++p) host[id[p]] = tmp[p];
    return QP_OK;
}

int host_ids(const Bunch *b, std::vector<int> &id) {
    id.resize(b->np);
    QP_HIP_CHECK(hipStreamSynchronize(b->stream()));
    if (b->np) QP_HIP_CHECK(hipMemcpy(id.data(), b->cur.id, sizeof(int) * b->np, hipMemcpyDeviceToHost));
    return QP_OK;
}

}  // namespace

extern "C" {

int qp_bunch_create(qp_ctx ctx, int nx, int ny, int nz, double dx, double dy, double dz, double field_schwinger,
                    qp_bunch *out) {
    BARG_CHECK(ctx && out, "qp_bunch_create: null argument");
    BARG_CHECK(nx >= 2 && ny >= 2 && nz >= 2, "qp_bunch_create: need nx, ny, nz >= 2");
    BARG_CHECK(dx > 0 && dy > 0 && dz > 0 && field_schwinger > 0, "qp_bunch_create: steps and field must be > 0");
    *out = nullptr;
    qp_bunch h = new qp_bunch_s();
    Bunch *b = &h->b;
    b->ctx = ctx;
    b->nx = nx;
    b->ny = ny;
    b->nz = nz;
    b->dx = dx;
    b->dy = dy;
    b->dz = dz;
    b->fs = field_schwinger;
    for (BunchDev *d : {&b->cur, &b->tmp}) {
        d->nx = nx;
        d->n1 = ny;
        d->n2 = nz;
        d->dx = dx;
        d->dy = dy;
        d->dz = dz;
        d->fs = field_schwinger;
    }
    auto fail = [&](int st) {
        qp_bunch_destroy(h);
        return st;
    };
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(QP_ERR_HIP);
    if (hipMalloc(&b->start, sizeof(int) * (2 * nx + 1)) != hipSuccess ||
        hipMalloc(&b->count, sizeof(unsigned) * 2 * nx) != hipSuccess ||
        hipMalloc(&b->cursor, sizeof(unsigned) * 2 * nx) != hipSuccess ||
        hipMalloc(&b->planes, sizeof(double) * 2 * b->nn()) != hipSuccess) {
        set_error("qp_bunch_create: device allocation failed");
        return fail(QP_ERR_ALLOC);
    }
    hipStream_t s = b->stream();
    if (hipMemsetAsync(b->start, 0, sizeof(int) * (2 * nx + 1), s) != hipSuccess ||
        hipMemsetAsync(b->planes, 0, sizeof(double) * 2 * b->nn(), s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(QP_ERR_HIP);
    b->cur.start = b->start;
    b->tmp.start = b->start;
    *out = h;
    return QP_OK;
}

int qp_bunch_destroy(qp_bunch h) {
    if (!h) return QP_OK;
    Bunch *b = &h->b;
    if (b->attached) {
        set_error("qp_bunch_destroy: a wake still streams this bunch (detach or destroy it first)");
        return QP_ERR_STATE;
    }
    if (b->ctx) (void)hipStreamSynchronize(b->stream());
    free_set(b->cur);
    free_set(b->tmp);
    for (void *p : {(void *)b->start, (void *)b->count, (void *)b->cursor, (void *)b->key, (void *)b->planes})
        if (p) (void)hipFree(p);
    delete h;
    return QP_OK;
}

int qp_bunch_set_particles(qp_bunch h, size_t np, const double *x, const double *y, const double *z,
                           const double *px, const double *py, const double *pz, const double *gamma,
                           const double *q, const double *cmr, const int *photon) {
    BARG_CHECK(h, "null bunch");
    BARG_CHECK(np == 0 || (x && y && z && px && py && pz && gamma && q && cmr && photon),
               "qp_bunch_set_particles: every array is required");
    BARG_CHECK(np < (size_t)INT_MAX / 2, "qp_bunch_set_particles: too many particles");
    Bunch *b = &h->b;
    hipStream_t s = b->stream();
    QP_HIP_CHECK(hipStreamSynchronize(s));
    if (np > b->cap) {
        free_set(b->cur);
        free_set(b->tmp);
        if (b->key) (void)hipFree(b->key);
        b->key = nullptr;
        b->cap = 0;
        QP_TRY(alloc_set(b->cur, np));
        QP_TRY(alloc_set(b->tmp, np));
        QP_HIP_CHECK(hipMalloc(&b->key, sizeof(int) * np));
        b->cap = np;
        ++b->epoch;
    }
    b->np = np;
    b->cur.np = b->tmp.np = (int)np;
    if (np == 0) {
        QP_HIP_CHECK(hipMemsetAsync(b->start, 0, sizeof(int) * (2 * b->nx + 1), s));
        QP_HIP_CHECK(hipStreamSynchronize(s));
        return QP_OK;
    }
    BunchDev &T = b->tmp;
    const double *src[] = {x, y, z, px, py, pz, gamma, q, cmr};
    double *dst[] = {T.x, T.y, T.z, T.px, T.py, T.pz, T.gamma, T.q, T.cmr};
    for (int a = 0; a < 9; ++a) QP_HIP_CHECK(hipMemcpy(dst[a], src[a], sizeof(double) * np, hipMemcpyHostToDevice));
    QP_HIP_CHECK(hipMemcpy(T.photon, photon, sizeof(int) * np, hipMemcpyHostToDevice));
    std::vector<int> ids(np);
    for (size_t p = 0; p < np; ++p) ids[p] = (int)p;
    QP_HIP_CHECK(hipMemcpy(T.id, ids.data(), sizeof(int) * np, hipMemcpyHostToDevice));
    for (double *f : {T.ex, T.ey, T.ez, T.by, T.bz, T.chi}) QP_HIP_CHECK(hipMemsetAsync(f, 0, sizeof(double) * np, s));
    QP_HIP_CHECK(hipMemsetAsync(b->count, 0, sizeof(unsigned) * 2 * b->nx, s));
    BLAUNCH(k_bunch_key, bblk(np), kBT, T, b->key, (unsigned *)b->count);
    QP_TRY(bucket_sort(b, T, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

int qp_bunch_get_particles(qp_bunch h, size_t np, double *x, double *y, double *z, double *px, double *py,
                           double *pz, double *gamma) {
    BARG_CHECK(h, "null bunch");
    Bunch *b = &h->b;
    BARG_CHECK(np == b->np, "qp_bunch_get_particles: count mismatch");
    std::vector<int> id;
    QP_TRY(host_ids(b, id));
    const double *src[] = {b->cur.x, b->cur.y, b->cur.z, b->cur.px, b->cur.py, b->cur.pz, b->cur.gamma};
    double *dst[] = {x, y, z, px, py, pz, gamma};
    for (int a = 0; a < 7; ++a) QP_TRY(unsort(b, src[a], dst[a], id));
    return QP_OK;
}

int qp_bunch_get_fields(qp_bunch h, size_t np, double *ex, double *ey, double *ez, double *by, double *bz,
                        double *chi) {
    BARG_CHECK(h, "null bunch");
    Bunch *b = &h->b;
    BARG_CHECK(np == b->np, "qp_bunch_get_fields: count mismatch");
    std::vector<int> id;
    QP_TRY(host_ids(b, id));
    const double *src[] = {b->cur.ex, b->cur.ey, b->cur.ez, b->cur.by, b->cur.bz, b->cur.chi};
    double *dst[] = {ex, ey, ez, by, bz, chi};
    for (int a = 0; a < 6; ++a) QP_TRY(unsort(b, src[a], dst[a], id));
    return QP_OK;
}

int qp_bunch_push(qp_bunch h, double dt) {
    BARG_CHECK(h, "null bunch");
    BARG_CHECK(std::isfinite(dt), "qp_bunch_push: dt must be finite");
    Bunch *b = &h->b;
    if (!b->np) return QP_OK;
    hipStream_t s = b->stream();
    QP_HIP_CHECK(hipMemsetAsync(b->count, 0, sizeof(unsigned) * 2 * b->nx, s));
    BLAUNCH(k_bunch_push, bblk(b->np), kBT, b->cur, b->tmp, dt, b->key, (unsigned *)b->count);
    QP_TRY(bucket_sort(b, b->tmp, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

int qp_bunch_deposit(qp_bunch h, double *rho_b, double *jx_b) {
    BARG_CHECK(h && rho_b && jx_b, "qp_bunch_deposit: null argument");
    Bunch *b = &h->b;
    hipStream_t s = b->stream();
    const size_t vol = (size_t)b->nx * b->nn();
    double *d3 = nullptr;
    QP_HIP_CHECK(hipMalloc(&d3, sizeof(double) * 2 * vol));
    struct Free {
        double *p;
        ~Free() { (void)hipFree(p); }
    } guard{d3};
    QP_HIP_CHECK(hipMemsetAsync(d3, 0, sizeof(double) * 2 * vol, s));
    if (b->np) BLAUNCH(k_bunch_dep3d, bblk(b->np), kBT, b->cur, d3, d3 + vol);
    QP_HIP_CHECK(hipStreamSynchronize(s));
    QP_HIP_CHECK(hipMemcpy(rho_b, d3, sizeof(double) * vol, hipMemcpyDeviceToHost));
    QP_HIP_CHECK(hipMemcpy(jx_b, d3 + vol, sizeof(double) * vol, hipMemcpyDeviceToHost));
    return QP_OK;
}

int qp_bunch_size(qp_bunch h, size_t *np) {
    BARG_CHECK(h && np, "null argument");
    *np = h->b.np;
    return QP_OK;
}

}  // extern "C"
