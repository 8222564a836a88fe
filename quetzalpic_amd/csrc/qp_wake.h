// qp_wake.h -- device state of one wake (host side).
#pragma once

#include <vector>

#include "qp_bunch.h"
#include "qp_envelope.h"
#include "qp_internal.h"

namespace qp {

// Gaussian laser <a^2> evaluated per slice (src/Config_reader.cpp:114-128,
// src/profiles.cpp:4-11): amp * exp(-(x-x0)^2/wx^2 - (y-y0)^2/wy^2 - (z-z0)^2/wz^2).
struct LaserGauss {
    double amp, wx, wy, wz, x0, y0, z0;
};

// Plain FP64 particle state: the fields of wake_particle_2d
// (src/containers.h:31-43) as separate device arrays.
struct PartFP64 {
    double *y = nullptr, *z = nullptr, *py = nullptr, *pz = nullptr, *n = nullptr, *gamma = nullptr,
           *px = nullptr, *ym = nullptr, *zm = nullptr, *pyn = nullptr, *pzn = nullptr;
};

// Host data of one member of a batch of wakes (a single wake is a batch of
// one): the initial particles (FP64, re-uploaded every pass) and the laser.
struct WakeMember {
    size_t np = 0;
    bool particles_set = false;
    std::vector<double> h_y, h_z, h_py, h_pz, h_w;
    bool laser_set = false;
    LaserGauss laser{};
};

struct Wake {
    qp_ctx_s *ctx = nullptr;
    qp_wake_params p{};
    Solver2d S;
    int nw = 1;                    // batch members (mixed path only when > 1)
    std::vector<WakeMember> mem;   // nw entries
    size_t cap = 0;  // allocated particle capacity
    bool particles_dirty = true;  // host arrays changed since the device initial state was built

    PartFP64 pf;
    double *init64[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // y, z, py, pz, n of a pass

    // grids: nw planes of ny*nz FP64 side by side each (member m at + m*ny*nz)
    double *g_base = nullptr;
    double *src = nullptr, *psi = nullptr, *psi_prev = nullptr, *psim = nullptr, *rho = nullptr,
           *jx = nullptr, *jy = nullptr, *jz = nullptr, *jyn = nullptr, *jzn = nullptr, *djy = nullptr,
           *djz = nullptr, *rho_ion = nullptr, *by = nullptr, *bz = nullptr, *chi = nullptr, *ex = nullptr,
           *ey = nullptr, *ez = nullptr, *src2 = nullptr, *asq2 = nullptr, *psi_b[2] = {nullptr, nullptr},
           *zero2 = nullptr, *bx = nullptr;
    // driver inputs (the 3D arrays only for a single wake)
    double *a_sqr3 = nullptr;            // nx*ny*nz or null
    LaserGauss *d_laser = nullptr;       // nw members' lasers on the device
    double *rho_b3 = nullptr, *jx_b3 = nullptr;  // nx*ny*nz or null

    Ctl *ctl = nullptr;                  // nw control blocks
    double *h_out = nullptr;             // pinned host staging, 12 planes
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // mixed path: particle tile (threads per block, LDS window nodes per grid),
    // from the tile autotuner on the first pass unless set by QP_TILE
    int tile_pt = 256, tile_cap = 640;
    bool tile_tuned = false;
    // mixed path, last pass, per member: particle blocks of deposit launches and
    // those whose deposit window overflowed the LDS capacity
    std::vector<long long> win_blocks, win_ovf;

    // mixed-path state (owned by qp_wake_mixed.hip)
    void *mixed = nullptr;

    // Attached laser envelope (or null): <a^2> comes from its |a|^2/2 and the
    // chi of every slice goes to its chi volume.
    Envelope *env = nullptr;
    double env_dtau = 0.0;   // > 0: each pass also advances the envelope by one step (after it)
    int env_chi_it = 1;
    // Attached driver bunch (or null): its sources are deposited per slice into
    // two planes (in place of rho_b3/jx_b3) and it gathers the fields per slice.
    Bunch *bunch = nullptr;
    unsigned bunch_epoch = 0;  // Bunch::epoch the captured slice steps saw
    // Bumped by every change of the driver inputs (a_sqr, laser, bunch sources,
    // envelope): captured slice graphs hold their pointers and flags.
    unsigned inputs_epoch = 0;

    // profiling hooks: when non-null, launches of `prof_class` are bracketed by events
    const char *prof_class = nullptr;
    std::vector<hipEvent_t> prof_ev;
    int prof_count = 0;

    size_t nn() const { return (size_t)p.ny * p.nz; }
    size_t np() const { return mem[0].np; }  // particles of a single wake
    bool any_laser() const {
        for (const WakeMember &m : mem)
            if (m.laser_set) return true;
        return false;
    }
    hipStream_t stream() const { return ctx->stream; }
    // 3D <a^2> the slice loop reads (null: Gaussian laser or none).
    double *asq3() const { return env ? env->asq3 : a_sqr3; }
    double *chi_store(int slice) const { return env ? env->chi3 + (size_t)slice * nn() : nullptr; }
    bool env_step() const { return env && env_dtau > 0; }
    // Bunch sources the slice loop reads: slice i at base + i * stride (stride 0
    // for the per-slice planes of an attached bunch).
    const double *rho_b_src() const { return bunch ? bunch->rho_plane() : rho_b3; }
    const double *jx_b_src() const { return bunch ? bunch->jx_plane() : jx_b3; }
    size_t bsrc_stride() const { return bunch ? 0 : nn(); }
};

int wake_alloc_common(Wake *w);
int wake_run_fp64(Wake *w, qp_slice_cb cb, void *user, int *stop_slice);
// Batch slice callback: member index, then the slice view.
using batch_cb = int (*)(int member, const qp_slice_view *v, void *user);
int wake_run_mixed(Wake *w, batch_cb cb, void *user, int *stop_slices);
int wake_upload_fp64(Wake *w);
int wake_upload_mixed(Wake *w);
int wake_download_fp64(Wake *w, double *y, double *z, double *py, double *pz);
int wake_download_mixed(Wake *w, int member, double *y, double *z, double *py, double *pz);
void wake_free_particles(Wake *w);
void wake_destroy_mixed(Wake *w);

// Fill the 2D <a^2> plane of slice i (from the 3D host upload or the laser).
int wake_asq_slice(Wake *w, int i, const double **asq);
// Pointers to slice i of the bunch sources (null when absent).
const double *wake_rho_b(Wake *w, int i);
const double *wake_jx_b(Wake *w, int i);

// Profiling bracket helpers (no-ops unless a profile pass is active).
void prof_begin(Wake *w, const char *cls);
void prof_end(Wake *w, const char *cls);

// Stage the output planes of slice i to host and call cb (bx may be null).
int deliver_slice(Wake *w, qp_slice_cb cb, void *user, int slice, int kind, const double *psi,
                  const double *rho, const double *jx, const double *jy, const double *jz, const double *ex,
                  const double *ey, const double *ez, const double *by, const double *bz, const double *chi,
                  const double *bx = nullptr);

// A group of batches with possibly different transverse grids, advanced in
// lockstep with one launch per phase of the slice step over all of them.
// The batches stay owned by their handles.
struct Group {
    qp_ctx_s *ctx = nullptr;
    std::vector<Wake *> b;
    void *dev = nullptr;   // launch tables and graphs (owned by qp_wake_mixed.hip)
    int tile_pt = 256, tile_cap = 640;   // particle tile of every batch in the group's launches
    bool tile_tuned = false;
    bool fork_families = true;   // the two transform families' solver launches on two streams
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};
// Particle tiles the mixed kernels are instantiated for.
bool mixed_tile_valid(int pt, int cap);
std::string mixed_tiles_json();
// Group slice callback: batch index, member index, then the slice view.
using group_cb = int (*)(int batch, int member, const qp_slice_view *v, void *user);
int group_run_mixed(Group *g, group_cb cb, void *user, int *stop_slices);
void group_destroy_mixed(Group *g);

}  // namespace qp

struct qp_wake_s {
    qp::Wake w;
};

struct qp_group_s {
    qp::Group g;
};
