/*
 * quetzal_hip.h -- C-ABI of the MI355X-native quasistatic slice loop.
 *
 * Plain C, plain pointers and sizes.  Every entry point returns an int status
 * (QP_OK = 0, QP_STOPPED = 1 for the reference's soft stop, < 0 errors) and
 * leaves a message for qp_last_error() (thread-local).  Host buffers passed in
 * are borrowed for the duration of the call; device memory is owned by the
 * handle.  One handle per host thread.
 *
 * What each entry point replaces in QuetzalPIC (file:line of the reference):
 *   qp_fourier2d_*   class Fourier2d (src/Fourier2d.h:4-19, src/Fourier2d.cpp:3-43)
 *   qp_solve2d*      System_3d::solve_poisson_equation (src/System_3d.cpp:686-706)
 *   qp_wake_*        System_3d::solve_wakefield (src/System_3d.cpp:238-627) with
 *                    its particle state (src/containers.h:31-43,
 *                    src/System_3d.cpp:724-748) and output_step
 *                    (src/System_3d.cpp:629-666)
 */
#ifndef QUETZAL_HIP_H
#define QUETZAL_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QP_OK 0
#define QP_STOPPED 1          /* stop rule fired (src/System_3d.cpp:541-591) */
#define QP_ERR_ARG (-1)
#define QP_ERR_HIP (-2)
#define QP_ERR_STATE (-3)
#define QP_ERR_ALLOC (-4)

/* Arithmetic of the slice loop. */
#define QP_PREC_FP64 0   /* plain path: FP64 SoA particles, one kernel per reference loop */
#define QP_PREC_MIXED 1  /* fused path: u32 cell + u32 fixed-point in-cell offsets, FP64
                            momenta, weights and predicted momenta p*;
                            FP64 arithmetic, grids and solves; hipGraph slice step */

typedef struct qp_ctx_s *qp_ctx;
typedef struct qp_fourier2d_s *qp_fourier2d;
typedef struct qp_wake_s *qp_wake;

/* Thread-local message of the last failing call ("" if none). */
const char *qp_last_error(void);
/* Library version string. */
const char *qp_version(void);

/* Device context: HIP device index and one stream (default priority). */
int qp_ctx_create(int device, qp_ctx *out);
int qp_ctx_destroy(qp_ctx ctx);
int qp_ctx_synchronize(qp_ctx ctx);

/* ---- Fourier2d (src/Fourier2d.h:4-19) --------------------------------------
 * n1 x n2 real <-> n1 x (n2/2+1) complex (interleaved re,im), unnormalised,
 * FFTW r2c/c2r conventions (src/Fourier2d.cpp:33-34). */
int qp_fourier2d_create(qp_ctx ctx, int n1, int n2, qp_fourier2d *out);
int qp_fourier2d_destroy(qp_fourier2d f);
/* Fourier2d::forward_transform (src/Fourier2d.cpp:37-39): host in -> host out. */
int qp_fourier2d_forward(qp_fourier2d f, const double *in_real, double *out_complex);
/* Fourier2d::backward_transform (src/Fourier2d.cpp:41-43): host in -> host out. */
int qp_fourier2d_backward(qp_fourier2d f, const double *in_complex, double *out_real);

/* ---- solve_poisson_equation (src/System_3d.cpp:686-706) --------------------
 * dst = c2r( r2c(src) with (0,0) := 0 and each other mode divided by
 *            -D + (2cos(2 pi j/n1)-2)/d1^2 + (2cos(2 pi k/n2)-2)/d2^2 ),
 * unnormalised like the reference (the caller divides by n1*n2).
 * src/dst may alias.  Host pointers. */
int qp_solve2d(qp_fourier2d f, double d1, double d2, double D, const double *src, double *dst);
/* Same with device pointers, enqueued on the context stream (no sync). */
int qp_solve2d_device(qp_fourier2d f, double d1, double d2, double D, const double *src_dev,
                      double *dst_dev);

/* Allocation-free variant of qp_solve2d_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Same solve, enqueued on ctx's
 * stream; src_dev/dst_dev are device buffers of n1*n2 doubles (may alias).
 * The half spectrum, twiddle tables and eigenvalue terms live in one fixed
 * device array per GPU (8.07 MiB, held until the process ends), so n1 and n2
 * must lie in [1, 1024] (QP_ERR_ARG beyond).  It calls no hipMalloc/hipFree.
 * Calls that overlap in time on one GPU share that array and race. */
int qp_solve2d_device_noalloc(qp_ctx ctx, int n1, int n2, double d1, double d2, double D,
                              const double *src_dev, double *dst_dev);

/* ---- One wake: System_3d::solve_wakefield (src/System_3d.cpp:238-627) ---- */
typedef struct {
    int nx, ny, nz;          /* slices and transverse nodes, n = (int)(l/d) (src/System_3d.cpp:87-89) */
    double dx, dy, dz;       /* steps d = l/n (src/System_3d.cpp:93-95) */
    double ly, lz;           /* periodic transverse box */
    int K;                   /* magnetic_field_iterations (>= 0) */
    double D;                /* magnetic_field_D, reference 10 (src/System_3d.h:59) */
    double psi_threshold;    /* reference 1e-4 (src/System_3d.h:57) */
    double b_threshold;      /* THRESHOLD_B, reference 100 (src/System_3d.cpp:326) */
    int precision;           /* QP_PREC_FP64 | QP_PREC_MIXED */
    int use_graph;           /* mixed path: capture the slice step in a hipGraph (1) or not (0) */
    int resort_every;        /* mixed path: re-bin particles by cell tile every N slices (0 = never) */
    int longitudinal_b;      /* mixed path, extension with no reference counterpart: 1 = also solve
                                lap_perp B_x = dz j_y - dy j_z per slice (output plane bx) */
    int async_output;        /* mixed path: 1 = stream the per-slice output: the planes are
                                snapshotted on the device, copied on a second stream into pinned
                                host slots and handed to the callback by a writer thread (in
                                slice order, during the run call) while later slices compute;
                                the slice step stays a hipGraph.  0 = the callback runs inline. */
    int slice_limit;         /* no reference counterpart: > 0 runs only the first slice_limit
                                slices of the pass (nx-1 down to nx-slice_limit), each exactly as
                                in the whole pass; the final output_step(0), the bunch gather of
                                slice 0 and an enqueued envelope step are then not run.  0 = the
                                whole pass.  For bounded checks and timings of large grids. */
} qp_wake_params;

/* Fields of one slice handed to the per-slice callback (FP64, host memory,
 * valid during the callback only; ny*nz row-major [j][k]).
 * kind QP_SLICE_NORMAL: output_step(i) inside the loop (src/System_3d.cpp:610);
 * kind QP_SLICE_STOPPED: the slice where the stop rule fired -- only by, bz, chi
 *   are meaningful (the reference breaks before output_step, :582-591);
 * kind QP_SLICE_FINAL: output_step(0) after the loop (:622). */
#define QP_SLICE_NORMAL 0
#define QP_SLICE_STOPPED 1
#define QP_SLICE_FINAL 2
typedef struct {
    int slice;
    int kind;
    const double *psi, *rho, *jx, *jy, *jz;   /* 2D arrays written by write_slice */
    const double *ex, *ey, *ez, *by, *bz, *chi;  /* slice i of the 3D arrays */
    const double *bx;        /* longitudinal B at the cell corners (dy/2, dz/2), from the j_y, j_z
                                of this view; NULL unless longitudinal_b */
} qp_slice_view;
typedef int (*qp_slice_cb)(const qp_slice_view *view, void *user);

int qp_wake_create(qp_ctx ctx, const qp_wake_params *params, qp_wake *out);
int qp_wake_destroy(qp_wake w);
/* Plasma macro-particles, FP64 SoA (wake_particle_2d y, z, py, pz, n;
 * src/containers.h:31-43).  py/pz may be NULL (cold plasma).  The library
 * converts them to its own device layout.  Called before every pass (the
 * reference re-initialises its particles each pass, src/System_3d.cpp:262). */
int qp_wake_set_particles(qp_wake w, size_t np, const double *y, const double *z,
                          const double *py, const double *pz, const double *weight);
/* Laser <a^2> on the nodes, nx*ny*nz host array (a_sqr, src/System_3d.cpp:805-816). */
int qp_wake_set_a_sqr(qp_wake w, const double *a_sqr);
/* Gaussian laser evaluated on the device per slice: 0.5*a0^2 exp(-sum (x-x0)^2/w^2)
 * with w = full 1/e^2 width / sqrt(8) (src/Config_reader.cpp:114-128). */
int qp_wake_set_laser_gaussian(qp_wake w, double a0, double xsigma, double ysigma, double zsigma,
                               double x0, double y0, double z0);
/* Driver-bunch sources rho_bunch, jx_bunch (nx*ny*nz host arrays, either may be
 * NULL = 0) (src/System_3d.cpp:345,384-385,449). */
int qp_wake_set_bunch_sources(qp_wake w, const double *rho_b, const double *jx_b);
/* Run one pass of the slice loop.  cb may be NULL (no per-slice host copies).
 * Returns QP_OK or QP_STOPPED (*stop_slice = slice index, else -1). */
int qp_wake_run(qp_wake w, qp_slice_cb cb, void *user, int *stop_slice);
/* Device time of the last qp_wake_run in milliseconds (hipEvents around the pass). */
int qp_wake_last_time_ms(qp_wake w, double *ms);
/* Instrumented pass: same work as qp_wake_run with cb = NULL, with hipEvents
 * around every launch of the named kernel class; returns launches and the mean
 * device duration in microseconds.  Classes: "push" (particle phases),
 * "solve" (all transform launches), "deposit_push_B" etc. (see docs). */
int qp_wake_profile_kernel(qp_wake w, const char *kernel_class, int *launches, double *mean_us);
/* Mixed path, last pass, one member (member 0 for a single wake): particle
 * blocks of the deposit launches (A, B, C, S phases) and those whose deposit
 * window exceeded the block's LDS capacity and deposited with global atomics
 * instead.  Particles are stored cell-sorted at the start of every pass;
 * resort_every re-bins them during it.  No reference counterpart. */
int qp_wake_window_stats(qp_wake w, int member, long long *overflow_blocks, long long *deposit_blocks);
/* Particle tile of the mixed path: threads (= particles) per block of the
 * push/deposit launches and LDS window nodes per deposited grid.  By default
 * the tile autotuner picks one per wake on its first pass (the first slices
 * with each instantiated tile, fastest kept; cached per grid and particle
 * count; QP_TILE=pt,cap in the environment forces one); this call forces a
 * tile, and (0, 0) lets the tuner pick again on the next pass.  The list of
 * instantiated tiles is in qp_wake_describe.  No reference counterpart. */
int qp_wake_set_tile(qp_wake w, int threads_per_block, int window_nodes);
/* JSON description of the launch shapes the library picked for this wake
 * (particle tile and the list of instantiated tiles; FFT rows/columns per
 * block of the generic kernels). */
int qp_wake_describe(qp_wake w, char *buf, size_t len);
/* Copy the current particle state back (FP64 SoA; any pointer may be NULL). */
int qp_wake_get_particles(qp_wake w, size_t np, double *y, double *z, double *py, double *pz);

/* ---- Batches of independent wakes (no reference counterpart) --------------
 * The reference runs one System_3d per process; its callers that sweep many
 * plasma / laser settings run many wakes.  A batch holds nwakes members with
 * the geometry and iteration settings of *params (QP_PREC_MIXED only).  The
 * members are advanced in lockstep, and every launch carries all of them.  Each
 * member has its own particles, Gaussian laser, fields and stop rule, and
 * gives the same result as the member run alone in its own handle.  The
 * single-wake calls above act on member 0; qp_wake_run requires nwakes == 1. */
int qp_wake_create_batch(qp_ctx ctx, const qp_wake_params *params, int nwakes, qp_wake *out);
int qp_wake_batch_size(qp_wake w, int *nwakes);
int qp_wake_set_particles_at(qp_wake w, int member, size_t np, const double *y, const double *z,
                             const double *py, const double *pz, const double *weight);
int qp_wake_set_laser_gaussian_at(qp_wake w, int member, double a0, double xsigma, double ysigma,
                                  double zsigma, double x0, double y0, double z0);
/* Per-slice callback of a batch: member index, then the view (as qp_slice_cb). */
typedef int (*qp_batch_cb)(int member, const qp_slice_view *view, void *user);
/* One pass for every member; stop_slices (nwakes entries, may be NULL) gets each
 * member's stop slice or -1.  Returns QP_STOPPED if any member stopped. */
int qp_wake_run_batch(qp_wake w, qp_batch_cb cb, void *user, int *stop_slices);
int qp_wake_get_particles_at(qp_wake w, int member, size_t np, double *y, double *z, double *py,
                             double *pz);

/* ---- Groups of batches with uneven grids (no reference counterpart) --------
 * A group advances several batches (each from qp_wake_create_batch, each with
 * its own transverse grid and members) in lockstep: every phase of the slice
 * step -- commit/advance, predictor, current deposits, and each pass of every
 * periodic solve -- is one launch over all members of all batches, and the
 * generic slice step is captured in a hipGraph.  Members give the same result
 * as their batch run alone.  The batches must be mixed-precision handles of
 * the group's context with equal nx, K and slice_limit, Gaussian lasers only
 * (no envelope, bunch, 3D a_sqr or bunch sources), no async_output,
 * resort_every or longitudinal_b, and grid sides 2^p in [16, 256] or 48, 96,
 * 192.  The group borrows the batch handles: destroy it before them.  One
 * group per host thread. */
typedef struct qp_group_s *qp_group;
int qp_group_create(qp_ctx ctx, int nbatches, const qp_wake *batches, qp_group *out);
int qp_group_destroy(qp_group g);
/* Per-slice callback of a group: batch index, member index, then the view. */
typedef int (*qp_group_cb)(int batch, int member, const qp_slice_view *view, void *user);
/* One pass of every member of every batch; stop_slices (sum of the batch sizes
 * entries, batch by batch; may be NULL) gets each member's stop slice or -1.
 * Returns QP_STOPPED if any member stopped. */
int qp_group_run(qp_group g, qp_group_cb cb, void *user, int *stop_slices);
/* Device time of the last qp_group_run in milliseconds. */
int qp_group_last_time_ms(qp_group g, double *ms);
/* Particle tile of every launch of the group (as qp_wake_set_tile; by default
 * the group's first run tunes one for all its batches together). */
int qp_group_set_tile(qp_group g, int threads_per_block, int window_nodes);
/* Solver passes of a group holding both transform families (sides 2^p and
 * 3 * 2^p): nonzero (the default) runs the two families' launches on two
 * streams at once (a fork and join inside the slice step); 0 runs them in
 * sequence on the context's stream, which is faster when several groups run
 * concurrently on one device.  Results are the same either way. */
int qp_group_set_fork(qp_group g, int fork_families);

/* ---- Laser envelope (no reference counterpart) ------------------------------
 * The reference keeps <a^2> fixed for the whole run (src/System_3d.cpp:805-816
 * is its only writer) and deposits chi (src/System_3d.cpp:396,526) without
 * reading it.  A qp_envelope holds the complex envelope a of the laser vector
 * potential Re[a exp(i k0 (x - t))] on an nx x ny x nz grid (slice i at
 * x = i dx, x toward the front; arrays [i][j][k] like a_sqr), k0 in units of
 * the plasma wavenumber, and advances it in the time tau of the moving window:
 *     lap_perp a + 2 d_tau (d_x + i k0) a = chi a
 * (Crank-Nicolson in tau, periodic 5-point lap_perp, d_x differenced toward
 * the front, slices swept from the front; chi minus its midrange on the slice
 * couples the transverse modes in the implicit half and is solved by
 * chi_iterations fixed-point corrections per slice, which contract for any
 * dtau; 0 = that part of chi a at the old time only).
 * Attached to a wake, the wake reads <a^2> = |a|^2/2 in place of its a_sqr or
 * laser and stores the chi of every slice it computes into the envelope
 * (slices after a stop keep their previous chi).  Single wakes only; the
 * envelope must share the wake's context and grid, and outlive the attachment. */
typedef struct qp_envelope_s *qp_envelope;
int qp_envelope_create(qp_ctx ctx, int nx, int ny, int nz, double dx, double dy, double dz, double k0,
                       qp_envelope *out);
/* Fails with QP_ERR_STATE while a wake is attached. */
int qp_envelope_destroy(qp_envelope e);
/* Host arrays nx*ny*nz; im may be NULL (real envelope). */
int qp_envelope_set(qp_envelope e, const double *re, const double *im);
/* a = a0 exp(-4 sum (x-x0)^2/sigma^2): <a^2> equals qp_wake_set_laser_gaussian's. */
int qp_envelope_set_gaussian(qp_envelope e, double a0, double xsigma, double ysigma, double zsigma,
                             double x0, double y0, double z0);
int qp_envelope_get(qp_envelope e, double *re, double *im);
/* chi of every slice, nx*ny*nz (NULL = zero). */
int qp_envelope_set_chi(qp_envelope e, const double *chi);
int qp_envelope_get_chi(qp_envelope e, double *chi);
/* <a^2> = |a|^2/2 as the attached wake reads it. */
int qp_envelope_get_a_sqr(qp_envelope e, double *a_sqr);
/* steps steps of dtau (one sweep each, captured in a hipGraph when use_graph),
 * chi_iterations in [0, 16]; ms (may be NULL) gets the device time of the call. */
int qp_envelope_advance(qp_envelope e, double dtau, int steps, int chi_iterations, int use_graph,
                        double *ms);
/* Attach e to the wake (NULL detaches). */
int qp_wake_attach_envelope(qp_wake w, qp_envelope e);
/* dtau > 0: every qp_wake_run also advances the attached envelope by one step
 * of dtau with the chi of that pass, as qp_wake_run followed by
 * qp_envelope_advance(e, dtau, 1, chi_iterations, ...), enqueued on the wake's
 * stream behind the pass (one coupled time step per call, no host round trip
 * in between).  dtau = 0 turns it off. */
int qp_wake_set_envelope_step(qp_wake w, double dtau, int chi_iterations);

/* ---- Driver bunch on the device (System_3d's bunch, src/System_3d.cpp:843-946) --
 * Replaces the host-side bunch of run() (src/System_3d.cpp:199-236):
 * deposit_bunch_densities (:926-946), deposit_wakefield_to_bunch (:843-860)
 * and advance_bunch_particles (:862-924) without its QED photon emission.
 * Particles are FP64 SoA: position, momentum, gamma, q = species charge * n,
 * cmr = species charge-to-mass ratio, photon = species is_photon.  Attached to
 * a single wake, qp_wake_run streams the bunch: slice i's rho_b, jx_b are
 * deposited into two planes right before the slice (in place of the 3D
 * arrays of qp_wake_set_bunch_sources) and, once slice i's fields are final,
 * the particles whose x-cell is i gather them with those of slice i + 1
 * (slice 0 after output_step(0)).  Particles whose x-cell is outside
 * [0, nx - 1) neither deposit nor gather (the reference's x cut,
 * src/array_utils.cpp:299-301).  Getters return particles in upload order. */
typedef struct qp_bunch_s *qp_bunch;
/* field_schwinger: the Schwinger field in plasma units (the chi normalisation). */
int qp_bunch_create(qp_ctx ctx, int nx, int ny, int nz, double dx, double dy, double dz,
                    double field_schwinger, qp_bunch *out);
/* Fails with QP_ERR_STATE while a wake is attached. */
int qp_bunch_destroy(qp_bunch b);
int qp_bunch_set_particles(qp_bunch b, size_t np, const double *x, const double *y, const double *z,
                           const double *px, const double *py, const double *pz, const double *gamma,
                           const double *q, const double *cmr, const int *photon);
int qp_bunch_size(qp_bunch b, size_t *np);
int qp_bunch_get_particles(qp_bunch b, size_t np, double *x, double *y, double *z, double *px,
                           double *py, double *pz, double *gamma);
/* Fields and chi gathered by the last qp_wake_run (zero where not gathered). */
int qp_bunch_get_fields(qp_bunch b, size_t np, double *ex, double *ey, double *ez, double *by,
                        double *bz, double *chi);
/* advance_bunch_particles without QED emission, then re-bucketing by x-cell. */
int qp_bunch_push(qp_bunch b, double dt);
/* deposit_bunch_densities into nx*ny*nz host arrays (the 3D output). */
int qp_bunch_deposit(qp_bunch b, double *rho_b, double *jx_b);
/* Attach b to the wake (NULL detaches). */
int qp_wake_attach_bunch(qp_wake w, qp_bunch b);

#ifdef __cplusplus
}
#endif
#endif
