// qp_bunch.h -- device state of a driver bunch (host side) and the per-slice
// launches the wake loops make for an attached bunch.
//
// The reference deposits the whole bunch into 3D arrays before the pass
// (deposit_bunch_densities, src/System_3d.cpp:926-946) and gathers the 3D
// field arrays after it (deposit_wakefield_to_bunch, :843-860).  Here the
// particles are kept sorted by half x-cell floor(2x/dx), and the wake streams
// both: at the start of slice i the particles whose x-cell has slice i as a
// face deposit rho_b(i), jx_b(i) into two planes; once slice i's fields are
// final, the particles for which slice i is a face of their x-cell in a field
// array (e_x is staggered by -dx/2 in x) take that half of their trilinear
// gather.  No 3D array and no plane of an earlier slice is kept.
#pragma once

#include "qp_internal.h"

namespace qp {

// What the kernels see (by value).
struct BunchDev {
    double *x, *y, *z, *px, *py, *pz, *gamma, *q, *cmr;  // q = species charge * n
    int *photon, *id;                                      // id = index at upload
    double *ex, *ey, *ez, *by, *bz, *chi;                  // gathered by the last pass
    const int *start;  // 2 nx + 1 entries: half-cell bucket b holds [start[b], start[b+1])
    int np, nx, n1, n2;
    double dx, dy, dz, fs;  // fs: Schwinger field in plasma units
};

struct Bunch {
    qp_ctx_s *ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    double dx = 0, dy = 0, dz = 0, fs = 1;
    size_t np = 0, cap = 0;
    BunchDev cur{}, tmp{};       // sorted arrays and the re-bucketing scratch
    int *start = nullptr, *count = nullptr, *cursor = nullptr, *key = nullptr;
    double *planes = nullptr;    // rho_b, jx_b planes of the current slice
    int attached = 0;            // wakes streaming this bunch
    unsigned epoch = 0;          // bumped when the particle arrays move (captured slice steps hold them)

    size_t nn() const { return (size_t)ny * nz; }
    double *rho_plane() const { return planes; }
    double *jx_plane() const { return planes + nn(); }
    hipStream_t stream() const { return ctx->stream; }
};

// Launches for an attached bunch, enqueued on s.  slot >= 0: the slice index
// comes from ctl->slice_of[slot] (captured slice steps); otherwise `slice`.
int bunch_clear_fields(Bunch *b, hipStream_t s);
int bunch_slice_sources(Bunch *b, const Ctl *ctl, int slot, int slice, hipStream_t s);
int bunch_slice_gather(Bunch *b, const double *ex, const double *ey, const double *ez, const double *by,
                       const double *bz, const Ctl *ctl, int slot, int slice, hipStream_t s);

}  // namespace qp

struct qp_bunch_s {
    qp::Bunch b;
};
