/*
 * qp_oracle.h -- FP64 CPU oracle for the quasistatic xi-slice loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product path (the HIP library behind include/quetzal_hip.h)
 * never links, loads or calls it.
 *
 * A plain-C restatement of System_3d::solve_wakefield (QuetzalPIC,
 * src/System_3d.cpp:238-706) and of the CIC kernels it calls
 * (src/array_utils.cpp:50-288).  Every function cites the reference lines it
 * follows.  The reference itself cannot be built here (FFTW3 and cpptoml are
 * absent), so the FFTW r2c/c2r transforms it calls are restated with an own
 * mixed-radix FFT that has FFTW's unnormalised conventions.
 *
 * Built twice from the same source (oracle/Makefile):
 *   liboracle.so      -O2 -ffp-contract=off, no OpenMP: serial, ordered
 *                     deposits, deterministic -- the parity oracle;
 *   liboracle_omp.so  -Ofast -fopenmp: the reference's Release flags
 *                     (CMakeLists.txt:23) and its OpenMP loop structure with
 *                     atomic deposits -- the timed CPU baseline.
 *
 * Pinning (all in tests/test_oracle_*.py): FFT against numpy.fft on many n;
 * the periodic solve against single-mode closed forms, an independent
 * 5-point stencil residual and numpy's rfft2/irfft2; CIC deposit/gather by
 * charge conservation, adjointness, exact bilinear reproduction and wrap
 * cases; derivative gathers by exact linear/quadratic fields; the whole loop
 * by the zero-driver identity, 4-fold symmetry, the gamma - px = 1 + psi
 * invariant and the linear-wake Green's-function known answer
 * (notebooks/Test3D.ipynb cell 1, mirrored to today's slice direction).
 */
#ifndef QP_ORACLE_H
#define QP_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid and loop constants of one wake (System_3d members, src/System_3d.h:45-59). */
typedef struct {
    int nx, ny, nz;        /* slices, transverse nodes (src/System_3d.cpp:87-89) */
    double dx, dy, dz;     /* steps after d = l/n (src/System_3d.cpp:93-95) */
    double ly, lz;         /* transverse box (periodic wrap length) */
    int K;                 /* magnetic_field_iterations */
    double D;              /* magnetic_field_D = 10 (src/System_3d.h:59) */
    double psi_threshold;  /* 1e-4 (src/System_3d.h:57) */
    double b_threshold;    /* THRESHOLD_B = 100 (src/System_3d.cpp:326) */
} qpo_grid;

/* Per-slice outputs, each nx*ny*nz row-major [i][j][k] or NULL.
 * psi..jz are the 2D arrays written per slice by output_step
 * (src/System_3d.cpp:629-666); ex..chi are the reference's 3D arrays. */
typedef struct {
    double *psi, *rho, *jx, *jy, *jz;
    double *ex, *ey, *ez, *by, *bz, *chi;
} qpo_fields;

/* ---- FFT (restates FFTW's unnormalised conventions) ------------------- */
/* In-place complex DFT of length n, interleaved (re,im); sign -1 forward. */
void qpo_fft(int n, double *data, int sign);
/* r2c 2D: in[n1*n2] -> out[n1*(n2/2+1)] complex interleaved
 * (fftw_plan_dft_r2c_2d, src/Fourier2d.cpp:33). */
void qpo_rfft2(int n1, int n2, const double *in, double *out);
/* c2r 2D, unnormalised (fftw_plan_dft_c2r_2d, src/Fourier2d.cpp:34).
 * 'in' is not modified. */
void qpo_irfft2(int n1, int n2, const double *in, double *out);

/* System_3d::solve_poisson_equation(D), src/System_3d.cpp:686-706:
 * r2c, zero the (0,0) mode, divide by the 5-point eigenvalue, c2r.
 * Unnormalised: inout holds n1*n2 times the solution on return. */
void qpo_solve_poisson(int n1, int n2, double d1, double d2, double D, double *inout);
/* Precision study (test infrastructure only): non-zero rounds every solve's
 * input, spectrum and output to single precision; 0 restores FP64. */
void qpo_set_solve_f32(int on);
/* Precision study (test infrastructure only): non-zero rounds the predicted
 * momentum increment p* - p to single precision; 0 restores FP64. */
void qpo_set_pstar_f32(int on);

/* ---- CIC deposit / gather (src/array_utils.cpp) ----------------------- */
/* deposit(y,z,v,array2d&) with an origin, :85-118 (also :50-83 for 3D slices). */
void qpo_deposit(double y, double z, double v, double *a, int n1, int n2,
                 double d1, double d2, double o1, double o2);
/* deposit(y,z,v,double*,d,n): truncation, no origin, :120-142. */
void qpo_deposit_raw(double y, double z, double v, double *a, int n1, int n2,
                     double d1, double d2);
/* array_to_particle(y,z,array2d&), :167-188 (also :144-165). */
double qpo_gather(double y, double z, const double *a, int n1, int n2,
                  double d1, double d2, double o1, double o2);
/* array_yder_to_particle / array_zder_to_particle, :240-288 (and :190-238). */
double qpo_gather_yder(double y, double z, const double *a, int n1, int n2, double d1, double d2);
double qpo_gather_zder(double y, double z, const double *a, int n1, int n2, double d1, double d2);

/* normalize_potential, src/System_3d.cpp:30-47. */
void qpo_normalize_potential(double *psi, int n1, int n2);
/* increase_minimum(array2d&, value), src/System_3d.cpp:830-841. */
void qpo_increase_minimum(double *a, int n, double value);
/* normalize_coordinates, src/System_3d.cpp:668-684. */
void qpo_normalize_coordinates(double *y, double *z, double ly, double lz);

/* ---- Initial state ------------------------------------------------------ */
/* Plasma profile kinds: reference constant2d/powerlaw2d
 * (src/profiles.cpp:45-63, src/Config_reader.cpp:236-258), plus the
 * parabolic channel extension -(base + curv*r^2) (no reference counterpart). */
enum { QPO_PROFILE_CONSTANT = 0, QPO_PROFILE_POWERLAW = 1, QPO_PROFILE_CHANNEL = 2 };
typedef struct {
    int kind;
    double rho0;                 /* constant: value (default -1) */
    double power, radius, min_factor, max_factor, y0, z0; /* powerlaw (amplitude -1) */
    double base, curv;           /* channel: -(base + curv*r^2) about (y0,z0) */
} qpo_profile;
double qpo_profile_eval(const qpo_profile *p, double y, double z);

/* count_wake_particles / init_wake_particles, src/System_3d.cpp:708-748.
 * Lattice ((i+1/2)dy/ppcy, (j+1/2)dz/ppcz), weight profile/ppcy/ppcz, zero
 * momenta; returns the count (arrays may be NULL to count only). */
size_t qpo_init_lattice(int ny, int nz, double dy, double dz, int ppcy, int ppcz,
                        const qpo_profile *prof, double *y, double *z, double *w);

/* init_a_sqr with gaussian3d(0.5*a0^2, width/sqrt(8), r0)
 * (src/System_3d.cpp:805-816, src/Config_reader.cpp:114-128, src/profiles.cpp:4-11).
 * a_sqr is nx*ny*nz, a_sqr(i,j,k) = f(i*dx, j*dy, k*dz). */
void qpo_fill_a_sqr_gaussian(const qpo_grid *g, double a0, double xsigma, double ysigma,
                             double zsigma, double x0, double y0, double z0, double *a_sqr);
/* The same for slices i0 .. i0+nheld-1 only, slice i at index i - i0. */
void qpo_fill_a_sqr_gaussian_window(const qpo_grid *g, int i0, int nheld, double a0, double xsigma,
                                    double ysigma, double zsigma, double x0, double y0, double z0,
                                    double *a_sqr);

/* ---- The slice loop ------------------------------------------------------ */
/* System_3d::solve_wakefield, src/System_3d.cpp:238-627, for one pass with
 * the given initial particles (the reference re-initialises them each pass,
 * :262).  a_sqr, rho_b, jx_b are nx*ny*nz (rho_b/jx_b may be NULL = 0).
 * Returns 0, or 1 when the stop rule fired (*stop_slice = slice index). */
int qpo_wake_run(const qpo_grid *g, size_t np, const double *y, const double *z,
                 const double *py, const double *pz, const double *w,
                 const double *a_sqr, const double *rho_b, const double *jx_b,
                 qpo_fields *out, int *stop_slice);

/* Same loop but only the first n_slices slices (for bounded CPU timing). */
int qpo_wake_run_partial(const qpo_grid *g, size_t np, const double *y, const double *z,
                         const double *py, const double *pz, const double *w,
                         const double *a_sqr, const double *rho_b, const double *jx_b,
                         qpo_fields *out, int *stop_slice, int n_slices);

/* The same loop over the first n_slices slices (nx-1 down to nx-n_slices) with
 * windowed 3D arrays: a_sqr, rho_b, jx_b and every output hold only those
 * slices, slice i at index i - (nx - n_slices).  For large grids, where the
 * full nx*ny*nz arrays would not fit in host memory. */
int qpo_wake_run_window(const qpo_grid *g, size_t np, const double *y, const double *z,
                        const double *py, const double *pz, const double *w,
                        const double *a_sqr, const double *rho_b, const double *jx_b,
                        qpo_fields *out, int *stop_slice, int n_slices);

/* ---- Driver bunch (src/System_3d.cpp:843-946) ----------------------------
 * 3D arrays are nx*ny*nz row-major [i][j][k] on the nodes (origin 0).
 * Index wrap: the reference computes j2 = (j1+1) % n before wrapping a
 * negative j1 once (src/array_utils.cpp:303-314); the oracle wraps both
 * indices fully, which is the same wherever the reference stays in bounds. */
/* deposit(x,y,z,v,array3d&), src/array_utils.cpp:290-338: nothing when the
 * x-cell i1 = floor(x/dx) is outside [0, nx-1). */
void qpo_deposit3d(const qpo_grid *g, double x, double y, double z, double v, double *a3);
/* array_to_particle(x,y,z,array3d&), src/array_utils.cpp:340-374, for an
 * array with origin (ox, oy, oz) (0 outside the x range). */
double qpo_gather3d(const qpo_grid *g, double ox, double oy, double oz, double x, double y, double z,
                    const double *a3);
/* deposit_bunch_densities, src/System_3d.cpp:926-946: rho_b = sum q S,
 * jx_b = sum q px/gamma S (both zeroed first); q = species charge * n. */
void qpo_bunch_deposit(const qpo_grid *g, size_t np, const double *x, const double *y, const double *z,
                       const double *px, const double *gamma, const double *q, double *rho_b, double *jx_b);
/* The same into arrays holding slices base .. base+nheld-1 only (slice i at
 * index i - base); deposits to other slices are dropped, the x cut of the
 * full grid still applies. */
void qpo_bunch_deposit_window(const qpo_grid *g, int base, int nheld, size_t np, const double *x,
                              const double *y, const double *z, const double *px, const double *gamma,
                              const double *q, double *rho_b, double *jx_b);
/* deposit_wakefield_to_bunch, src/System_3d.cpp:843-860: the five fields at
 * each particle and the QED parameter chi (field_schwinger in plasma units).
 * The field arrays carry the reference's staggered origins
 * (src/System_3d.cpp:179-183): ex (-dx/2, 0, 0), ey and bz (0, dy/2, 0),
 * ez and by (0, 0, dz/2). */
void qpo_bunch_fields(const qpo_grid *g, size_t np, const double *x, const double *y, const double *z,
                      const double *px, const double *py, const double *pz, const double *gamma,
                      const double *ex3, const double *ey3, const double *ez3, const double *by3,
                      const double *bz3, double field_schwinger, double *ex, double *ey, double *ez,
                      double *by, double *bz, double *chi);
/* advance_bunch_particles without QED emission, src/System_3d.cpp:862-890:
 * photons stream (gamma = |p|), charges take the explicit Lorentz step. */
void qpo_bunch_push(size_t np, double *x, double *y, double *z, double *px, double *py, double *pz,
                    double *gamma, const double *cmr, const int *photon, const double *ex,
                    const double *ey, const double *ez, const double *by, const double *bz, double dt);

/* Number of OpenMP threads the build uses (1 without OpenMP). */
int qpo_num_threads(void);

/* SplitMix64 counter-based stream: value for (seed, stream, index). */
unsigned long long qpo_splitmix64(unsigned long long seed, unsigned long long stream,
                                  unsigned long long index);

#ifdef __cplusplus
}
#endif
#endif
