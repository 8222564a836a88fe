/*
 * qp_oracle.c -- FP64 CPU oracle for the quasistatic xi-slice loop.
 * TEST INFRASTRUCTURE ONLY (see qp_oracle.h): loaded by tests/, smoke() and
 * bench.py's cpu_baseline leg as the checker, never by the product path.
 *
 * Restates QuetzalPIC's System_3d::solve_wakefield (src/System_3d.cpp:238-706)
 * and the CIC kernels of src/array_utils.cpp:50-288 in plain C.  Arithmetic
 * follows the reference's operation order so that the FP64 GPU path can be
 * compared at rounding level.  The only intentional difference: where the
 * reference indexes one past the grid (a coordinate whose y/d rounds to
 * exactly n; the reference's asserts are compiled out and it writes out of
 * bounds, src/array_utils.cpp:121-132), the oracle wraps the index to 0.
 */
#include "qp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int qpo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * FFT: recursive mixed-radix decimation in time, any n.
 * X[q*m + s] = sum_r w_p^{rq} (w_n^{rs} Y_r[s]), Y_r = DFT_m(x[r::p]).
 * Twiddles come from one table e^{-2 pi i t / N} evaluated in long double.
 * ---------------------------------------------------------------------- */
typedef struct {
    int n;
    int nfac;
    int fac[64];
    double *tw;      /* 2*n doubles: cos, -sin of 2 pi t / n */
    double *scratch; /* 2*n doubles per call site (not shared across threads) */
} fft_plan;

static void plan_init(fft_plan *p, int n) {
    p->n = n;
    p->nfac = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
        if (f * f > m && m > 1) { p->fac[p->nfac++] = m; m = 1; }
    }
    if (p->nfac == 0) p->fac[p->nfac++] = 1;
    p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
        long double ang = two_pi * (long double)t / (long double)n;
        p->tw[2 * t] = (double)cosl(ang);
        p->tw[2 * t + 1] = (double)(-sinl(ang));
    }
    p->scratch = NULL;
}

static void plan_free(fft_plan *p) {
    free(p->tw);
    p->tw = NULL;
}

/* w(t) = e^{sign * 2 pi i t / N} for t in [0, N). */
static inline void tw_get(const fft_plan *p, long t, int sign, double *re, double *im) {
    t %= p->n;
    *re = p->tw[2 * t];
    *im = sign < 0 ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
}

static void fft_rec(const fft_plan *pl, int n, const double *in, long is, double *out,
                    int sign, const int *fac, double *zbuf) {
    if (n == 1) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    const int p = fac[0];
    const int m = n / p;
    for (int r = 0; r < p; ++r)
        fft_rec(pl, m, in + 2 * (long)r * is, is * p, out + 2 * (long)r * m, sign, fac + 1, zbuf + 2 * p);
    const long step = pl->n / n;      /* w_n = w_N^{step} */
    const long pstep = pl->n / p;     /* w_p = w_N^{pstep} */
    double *z = zbuf;
    for (int s = 0; s < m; ++s) {
        for (int r = 0; r < p; ++r) {
            double ar = out[2 * ((long)r * m + s)], ai = out[2 * ((long)r * m + s) + 1];
            if (r == 0 || s == 0) {
                z[2 * r] = ar;
                z[2 * r + 1] = ai;
            } else {
                double wr, wi;
                tw_get(pl, step * r * s, sign, &wr, &wi);
                z[2 * r] = ar * wr - ai * wi;
                z[2 * r + 1] = ar * wi + ai * wr;
            }
        }
        if (p == 2) {
            out[2 * s] = z[0] + z[2];
            out[2 * s + 1] = z[1] + z[3];
            out[2 * (m + s)] = z[0] - z[2];
            out[2 * (m + s) + 1] = z[1] - z[3];
        } else if (p == 4) {
            /* w_4 = e^{sign i pi/2} = sign * i */
            double t0r = z[0] + z[4], t0i = z[1] + z[5];
            double t1r = z[0] - z[4], t1i = z[1] - z[5];
            double t2r = z[2] + z[6], t2i = z[3] + z[7];
            double t3r = z[2] - z[6], t3i = z[3] - z[7];
            /* (sign * i) * t3 */
            double u3r = -sign * t3i, u3i = sign * t3r;
            out[2 * s] = t0r + t2r;
            out[2 * s + 1] = t0i + t2i;
            out[2 * (m + s)] = t1r + u3r;
            out[2 * (m + s) + 1] = t1i + u3i;
            out[2 * (2 * m + s)] = t0r - t2r;
            out[2 * (2 * m + s) + 1] = t0i - t2i;
            out[2 * (3 * m + s)] = t1r - u3r;
            out[2 * (3 * m + s) + 1] = t1i - u3i;
        } else {
            for (int q = 0; q < p; ++q) {
                double accr = 0.0, acci = 0.0;
                for (int r = 0; r < p; ++r) {
                    double wr, wi;
                    tw_get(pl, pstep * (long)((r * (long)q) % p), sign, &wr, &wi);
                    accr += z[2 * r] * wr - z[2 * r + 1] * wi;
                    acci += z[2 * r] * wi + z[2 * r + 1] * wr;
                }
                out[2 * ((long)q * m + s)] = accr;
                out[2 * ((long)q * m + s) + 1] = acci;
            }
        }
    }
}

/* In-place transform using caller scratch (4*n + 2*sum(fac) doubles). */
static void fft_exec(const fft_plan *pl, double *data, int sign, double *work) {
    const int n = pl->n;
    double *out = work;
    double *zbuf = work + 2 * (size_t)n;
    fft_rec(pl, n, data, 1, out, sign, pl->fac, zbuf);
    memcpy(data, out, sizeof(double) * 2 * (size_t)n);
}

static size_t fft_work_size(const fft_plan *pl) {
    size_t s = 2 * (size_t)pl->n;
    for (int i = 0; i < pl->nfac; ++i) s += 2 * (size_t)pl->fac[i];
    return s + 16;
}

void qpo_fft(int n, double *data, int sign) {
    fft_plan pl;
    plan_init(&pl, n);
    double *work = (double *)malloc(sizeof(double) * fft_work_size(&pl));
    fft_exec(&pl, data, sign, work);
    free(work);
    plan_free(&pl);
}

/* ------------------------------------------------------------------------
 * 2D r2c / c2r with FFTW's layout: out[n1][n2/2+1] complex, unnormalised.
 * Rows are transformed two at a time packed as one complex sequence
 * z = x_a + i x_b; X_a[k] = (Z[k] + conj Z[-k])/2, X_b[k] = (Z[k] - conj Z[-k])/(2i).
 * ---------------------------------------------------------------------- */
typedef struct {
    int n1, n2, nh;
    fft_plan p1, p2;
} plan2d;

static void plan2d_init(plan2d *p, int n1, int n2) {
    p->n1 = n1;
    p->n2 = n2;
    p->nh = n2 / 2 + 1;
    plan_init(&p->p1, n1);
    plan_init(&p->p2, n2);
}

static void plan2d_free(plan2d *p) {
    plan_free(&p->p1);
    plan_free(&p->p2);
}

static void rfft2_exec(const plan2d *p, const double *in, double *out) {
    const int n1 = p->n1, n2 = p->n2, nh = p->nh;
    const int npairs = (n1 + 1) / 2;
#pragma omp parallel
    {
        double *z = (double *)malloc(sizeof(double) * 2 * (size_t)(n2 > n1 ? n2 : n1));
        size_t ws = fft_work_size(&p->p2) > fft_work_size(&p->p1) ? fft_work_size(&p->p2) : fft_work_size(&p->p1);
        double *work = (double *)malloc(sizeof(double) * ws);
#pragma omp for schedule(static)
        for (int a = 0; a < npairs; ++a) {
            const int ra = 2 * a, rb = 2 * a + 1;
            for (int t = 0; t < n2; ++t) {
                z[2 * t] = in[(size_t)ra * n2 + t];
                z[2 * t + 1] = rb < n1 ? in[(size_t)rb * n2 + t] : 0.0;
            }
            fft_exec(&p->p2, z, -1, work);
            for (int k = 0; k < nh; ++k) {
                const int kk = (n2 - k) % n2;
                double zr = z[2 * k], zi = z[2 * k + 1];
                double cr = z[2 * kk], ci = -z[2 * kk + 1];
                out[2 * ((size_t)ra * nh + k)] = 0.5 * (zr + cr);
                out[2 * ((size_t)ra * nh + k) + 1] = 0.5 * (zi + ci);
                if (rb < n1) {
                    /* (Z - conj Z[-k]) / (2i) = -i/2 (dr + i di) = (di/2, -dr/2) */
                    double dr = zr - cr, di = zi - ci;
                    out[2 * ((size_t)rb * nh + k)] = 0.5 * di;
                    out[2 * ((size_t)rb * nh + k) + 1] = -0.5 * dr;
                }
            }
        }
#pragma omp for schedule(static)
        for (int k = 0; k < nh; ++k) {
            for (int j = 0; j < n1; ++j) {
                z[2 * j] = out[2 * ((size_t)j * nh + k)];
                z[2 * j + 1] = out[2 * ((size_t)j * nh + k) + 1];
            }
            fft_exec(&p->p1, z, -1, work);
            for (int j = 0; j < n1; ++j) {
                out[2 * ((size_t)j * nh + k)] = z[2 * j];
                out[2 * ((size_t)j * nh + k) + 1] = z[2 * j + 1];
            }
        }
        free(z);
        free(work);
    }
}

/* c2r: inverse along n1 on the half spectrum, then a Hermitian inverse per
 * row (imaginary parts of the k=0 and, for even n2, k=n2/2 terms ignored,
 * as in FFTW's c2r).  'spec' is used as scratch and destroyed. */
static void irfft2_exec(const plan2d *p, double *spec, double *out) {
    const int n1 = p->n1, n2 = p->n2, nh = p->nh;
    const int npairs = (n1 + 1) / 2;
#pragma omp parallel
    {
        double *z = (double *)malloc(sizeof(double) * 2 * (size_t)(n2 > n1 ? n2 : n1));
        size_t ws = fft_work_size(&p->p2) > fft_work_size(&p->p1) ? fft_work_size(&p->p2) : fft_work_size(&p->p1);
        double *work = (double *)malloc(sizeof(double) * ws);
#pragma omp for schedule(static)
        for (int k = 0; k < nh; ++k) {
            for (int j = 0; j < n1; ++j) {
                z[2 * j] = spec[2 * ((size_t)j * nh + k)];
                z[2 * j + 1] = spec[2 * ((size_t)j * nh + k) + 1];
            }
            fft_exec(&p->p1, z, +1, work);
            for (int j = 0; j < n1; ++j) {
                spec[2 * ((size_t)j * nh + k)] = z[2 * j];
                spec[2 * ((size_t)j * nh + k) + 1] = z[2 * j + 1];
            }
        }
#pragma omp for schedule(static)
        for (int a = 0; a < npairs; ++a) {
            const int ra = 2 * a, rb = 2 * a + 1;
            for (int k = 0; k < n2; ++k) {
                double xr, xi, yr = 0.0, yi = 0.0;
                int kk = k;
                double conj = 1.0;
                if (k >= nh) { kk = n2 - k; conj = -1.0; }
                xr = spec[2 * ((size_t)ra * nh + kk)];
                xi = conj * spec[2 * ((size_t)ra * nh + kk) + 1];
                if (rb < n1) {
                    yr = spec[2 * ((size_t)rb * nh + kk)];
                    yi = conj * spec[2 * ((size_t)rb * nh + kk) + 1];
                }
                if (kk == 0 || (2 * kk == n2)) { xi = 0.0; yi = 0.0; }
                /* Z = X + i Y */
                z[2 * k] = xr - yi;
                z[2 * k + 1] = xi + yr;
            }
            fft_exec(&p->p2, z, +1, work);
            for (int t = 0; t < n2; ++t) {
                out[(size_t)ra * n2 + t] = z[2 * t];
                if (rb < n1) out[(size_t)rb * n2 + t] = z[2 * t + 1];
            }
        }
        free(z);
        free(work);
    }
}

void qpo_rfft2(int n1, int n2, const double *in, double *out) {
    plan2d p;
    plan2d_init(&p, n1, n2);
    rfft2_exec(&p, in, out);
    plan2d_free(&p);
}

void qpo_irfft2(int n1, int n2, const double *in, double *out) {
    plan2d p;
    plan2d_init(&p, n1, n2);
    size_t ns = 2 * (size_t)n1 * p.nh;
    double *tmp = (double *)malloc(sizeof(double) * ns);
    memcpy(tmp, in, sizeof(double) * ns);
    irfft2_exec(&p, tmp, out);
    free(tmp);
    plan2d_free(&p);
}

/* Precision study only (not a reference behaviour): with a non-zero value the
 * solve rounds its input, its spectrum before and after the eigenvalue
 * division, and its output to single precision, a lower bound on the
 * rounding of any transform that stores or accumulates in FP32 (an FP32 FFT,
 * or a split-BF16 dense product accumulated in FP32). */
static int g_solve_f32 = 0;
void qpo_set_solve_f32(int on) { g_solve_f32 = on; }
/* Precision study only: with a non-zero value the predictor's increment
 * p* - p is rounded to single precision (the mixed path's stored increment). */
static int g_pstar_f32 = 0;
void qpo_set_pstar_f32(int on) { g_pstar_f32 = on; }
static void round_f32(double *a, size_t n) {
    for (size_t i = 0; i < n; ++i) a[i] = (double)(float)a[i];
}

/* solve_poisson_equation(D), src/System_3d.cpp:686-706. */
static void solve_poisson_plan(const plan2d *p, double d1, double d2, double D, double *inout,
                               double *spec) {
    const int n1 = p->n1, n2 = p->n2, nh = p->nh;
    if (g_solve_f32) round_f32(inout, (size_t)n1 * n2);
    rfft2_exec(p, inout, spec);
    if (g_solve_f32) round_f32(spec, 2 * (size_t)n1 * nh);
    spec[0] = 0.0;
    spec[1] = 0.0;
#pragma omp parallel for
    for (int j = 0; j < n1; ++j) {
        for (int k = 0; k < nh; ++k) {
            if (k == 0 && j == 0) continue;
            double mult = -D + (2 * cos(2 * M_PI * j / n1) - 2) / d1 / d1 +
                          (2 * cos(2 * M_PI * k / n2) - 2) / d2 / d2;
            spec[2 * ((size_t)nh * j + k)] /= mult;
            spec[2 * ((size_t)nh * j + k) + 1] /= mult;
        }
    }
    if (g_solve_f32) round_f32(spec, 2 * (size_t)n1 * nh);
    irfft2_exec(p, spec, inout);
    if (g_solve_f32) round_f32(inout, (size_t)n1 * n2);
}

void qpo_solve_poisson(int n1, int n2, double d1, double d2, double D, double *inout) {
    plan2d p;
    plan2d_init(&p, n1, n2);
    double *spec = (double *)malloc(sizeof(double) * 2 * (size_t)n1 * p.nh);
    solve_poisson_plan(&p, d1, d2, D, inout, spec);
    free(spec);
    plan2d_free(&p);
}

/* ------------------------------------------------------------------------
 * CIC kernels, src/array_utils.cpp.
 * ---------------------------------------------------------------------- */
static inline void atomic_add(double *p, double v) {
#pragma omp atomic update
    *p += v;
}

/* deposit(y,z,value,array2d&), src/array_utils.cpp:85-118. */
void qpo_deposit(double y, double z, double v, double *a, int n1, int n2,
                 double d1, double d2, double o1, double o2) {
    int j1 = (int)floor((y - o1) / d1);
    int j2 = (j1 + 1) % n1;
    double yf = (y - o1) / d1 - j1;
    if (j1 < 0) j1 += n1;
    if (j1 >= n1) j1 -= n1;
    int k1 = (int)floor((z - o2) / d2);
    int k2 = (k1 + 1) % n2;
    double zf = (z - o2) / d2 - k1;
    if (k1 < 0) k1 += n2;
    if (k1 >= n2) k1 -= n2;
    atomic_add(&a[(size_t)n2 * j1 + k1], v * (1 - yf) * (1 - zf));
    atomic_add(&a[(size_t)n2 * j2 + k1], v * yf * (1 - zf));
    atomic_add(&a[(size_t)n2 * j1 + k2], v * (1 - yf) * zf);
    atomic_add(&a[(size_t)n2 * j2 + k2], v * yf * zf);
}

/* deposit(y,z,value,double*,d,n), src/array_utils.cpp:120-142. */
void qpo_deposit_raw(double y, double z, double v, double *a, int n1, int n2,
                     double d1, double d2) {
    int j1 = (int)(y / d1);
    int j2 = (j1 + 1) % n1;
    double yf = (y / d1) - j1;
    if (j1 >= n1) j1 -= n1;
    int k1 = (int)(z / d2);
    int k2 = (k1 + 1) % n2;
    double zf = (z / d2) - k1;
    if (k1 >= n2) k1 -= n2;
    atomic_add(&a[(size_t)n2 * j1 + k1], v * (1 - yf) * (1 - zf));
    atomic_add(&a[(size_t)n2 * j2 + k1], v * yf * (1 - zf));
    atomic_add(&a[(size_t)n2 * j1 + k2], v * (1 - yf) * zf);
    atomic_add(&a[(size_t)n2 * j2 + k2], v * yf * zf);
}

/* array_to_particle(y,z,array2d&), src/array_utils.cpp:167-188. */
double qpo_gather(double y, double z, const double *a, int n1, int n2,
                  double d1, double d2, double o1, double o2) {
    int j1 = (int)floor((y - o1) / d1);
    int j2 = (j1 + 1) % n1;
    double yf = (y - o1) / d1 - j1;
    if (j1 < 0) j1 += n1;
    if (j1 >= n1) j1 -= n1;
    int k1 = (int)floor((z - o2) / d2);
    int k2 = (k1 + 1) % n2;
    double zf = (z - o2) / d2 - k1;
    if (k1 < 0) k1 += n2;
    if (k1 >= n2) k1 -= n2;
    return a[(size_t)n2 * j1 + k1] * (1 - yf) * (1 - zf) + a[(size_t)n2 * j2 + k1] * yf * (1 - zf) +
           a[(size_t)n2 * j1 + k2] * (1 - yf) * zf + a[(size_t)n2 * j2 + k2] * yf * zf;
}

/* array_yder_to_particle(y,z,array2d&), src/array_utils.cpp:240-263. */
double qpo_gather_yder(double y, double z, const double *a, int n1, int n2, double d1, double d2) {
    int j1 = (int)floor(y / d1 - 0.5);
    double yf = (y / d1 - 0.5) - j1;
    if (j1 < 0) j1 += n1;
    int j2 = (j1 + 1) % n1;
    int j3 = (j1 + 2) % n1;
    int k1 = (int)(z / d2);
    int k2 = (k1 + 1) % n2;
    double zf = (z / d2) - k1;
    if (k1 >= n2) k1 -= n2;
    double g11 = (a[(size_t)n2 * j2 + k1] - a[(size_t)n2 * j1 + k1]) / d1;
    double g21 = (a[(size_t)n2 * j3 + k1] - a[(size_t)n2 * j2 + k1]) / d1;
    double g12 = (a[(size_t)n2 * j2 + k2] - a[(size_t)n2 * j1 + k2]) / d1;
    double g22 = (a[(size_t)n2 * j3 + k2] - a[(size_t)n2 * j2 + k2]) / d1;
    return g11 * (1 - yf) * (1 - zf) + g21 * yf * (1 - zf) + g12 * (1 - yf) * zf + g22 * yf * zf;
}

/* array_zder_to_particle(y,z,array2d&), src/array_utils.cpp:265-288. */
double qpo_gather_zder(double y, double z, const double *a, int n1, int n2, double d1, double d2) {
    int j1 = (int)(y / d1);
    int j2 = (j1 + 1) % n1;
    double yf = (y / d1) - j1;
    if (j1 >= n1) j1 -= n1;
    int k1 = (int)floor(z / d2 - 0.5);
    double zf = (z / d2 - 0.5) - k1;
    if (k1 < 0) k1 += n2;
    int k2 = (k1 + 1) % n2;
    int k3 = (k1 + 2) % n2;
    double g11 = (a[(size_t)n2 * j1 + k2] - a[(size_t)n2 * j1 + k1]) / d2;
    double g21 = (a[(size_t)n2 * j2 + k2] - a[(size_t)n2 * j2 + k1]) / d2;
    double g12 = (a[(size_t)n2 * j1 + k3] - a[(size_t)n2 * j1 + k2]) / d2;
    double g22 = (a[(size_t)n2 * j2 + k3] - a[(size_t)n2 * j2 + k2]) / d2;
    return g11 * (1 - yf) * (1 - zf) + g21 * yf * (1 - zf) + g12 * (1 - yf) * zf + g22 * yf * zf;
}

/* normalize_potential, src/System_3d.cpp:30-47: subtract the mean of the
 * n1+n2-1 border cells (column k=0, then row j=0 without its first cell). */
void qpo_normalize_potential(double *psi, int n1, int n2) {
    double border = 0.0;
    for (int i = 0; i < n1; ++i) border += psi[(size_t)n2 * i];
    for (int j = 1; j < n2; ++j) border += psi[j];
    border /= (n1 + n2 - 1);
#pragma omp parallel for
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) psi[(size_t)n2 * i + j] -= border;
}

/* increase_minimum(array2d&, value), src/System_3d.cpp:830-841. */
void qpo_increase_minimum(double *a, int n, double value) {
#pragma omp parallel for
    for (int i = 0; i < n; ++i)
        if (a[i] < value) a[i] = value;
}

/* normalize_coordinates, src/System_3d.cpp:668-684. */
void qpo_normalize_coordinates(double *y, double *z, double ly, double lz) {
    *y = fmod(*y, ly);
    if (*y < 0) *y += ly;
    *z = fmod(*z, lz);
    if (*z < 0) *z += lz;
}

/* ------------------------------------------------------------------------
 * Initial state.
 * ---------------------------------------------------------------------- */
double qpo_profile_eval(const qpo_profile *p, double y, double z) {
    switch (p->kind) {
    case QPO_PROFILE_POWERLAW: {
        /* powerlaw2d(power, radius, y0, z0, min, max, amplitude=-1), src/profiles.cpp:51-63 */
        double r = sqrt((y - p->y0) * (y - p->y0) + (z - p->z0) * (z - p->z0));
        double value = pow(r / p->radius, p->power);
        if (value > p->max_factor) value = p->max_factor;
        if (value < p->min_factor) value = p->min_factor;
        return -1.0 * value;
    }
    case QPO_PROFILE_CHANNEL: {
        double r2 = (y - p->y0) * (y - p->y0) + (z - p->z0) * (z - p->z0);
        return -(p->base + p->curv * r2);
    }
    default:
        return p->rho0; /* constant2d(rho0), src/profiles.cpp:45-49 */
    }
}

/* init_wake_particles, src/System_3d.cpp:724-748 (count: :708-722). */
size_t qpo_init_lattice(int ny, int nz, double dy, double dz, int ppcy, int ppcz,
                        const qpo_profile *prof, double *y, double *z, double *w) {
    size_t idx = 0;
    const double dppcy = ppcy, dppcz = ppcz;
    for (int i = 0; i < ppcy * ny; ++i) {
        for (int j = 0; j < ppcz * nz; ++j) {
            const double yy = (i + 0.5) * dy / ppcy;
            const double zz = (j + 0.5) * dz / ppcz;
            double value = qpo_profile_eval(prof, yy, zz);
            if (value != 0) {
                if (y) {
                    y[idx] = yy;
                    z[idx] = zz;
                    w[idx] = value / dppcy / dppcz;
                }
                idx++;
            }
        }
    }
    return idx;
}

void qpo_fill_a_sqr_gaussian_window(const qpo_grid *g, int i0, int nheld, double a0, double xsigma,
                                    double ysigma, double zsigma, double x0, double y0, double z0,
                                    double *a_sqr) {
    /* Config_reader::init_laser divides the full widths by sqrt(8) and uses
     * amplitude 0.5*a0^2 (src/Config_reader.cpp:124-128). */
    const double s8 = sqrt(8.0);
    const double wx = xsigma / s8, wy = ysigma / s8, wz = zsigma / s8;
    const double amp = 0.5 * a0 * a0;
    for (int i = i0; i < i0 + nheld; ++i)
        for (int j = 0; j < g->ny; ++j)
            for (int k = 0; k < g->nz; ++k) {
                double x = i * g->dx, y = j * g->dy, z = k * g->dz;
                a_sqr[((size_t)(i - i0) * g->ny + j) * g->nz + k] =
                    amp * exp(-(x - x0) * (x - x0) / wx / wx - (y - y0) * (y - y0) / wy / wy -
                              (z - z0) * (z - z0) / wz / wz);
            }
}

void qpo_fill_a_sqr_gaussian(const qpo_grid *g, double a0, double xsigma, double ysigma,
                             double zsigma, double x0, double y0, double z0, double *a_sqr) {
    qpo_fill_a_sqr_gaussian_window(g, 0, g->nx, a0, xsigma, ysigma, zsigma, x0, y0, z0, a_sqr);
}

unsigned long long qpo_splitmix64(unsigned long long seed, unsigned long long stream,
                                  unsigned long long index) {
    unsigned long long x = seed ^ (stream * 0x9E3779B97F4A7C15ULL) ^ (index * 0xD1B54A32D192ED03ULL);
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* ------------------------------------------------------------------------
 * The slice loop: System_3d::solve_wakefield, src/System_3d.cpp:238-627.
 * ---------------------------------------------------------------------- */
typedef struct {  /* wake_particle_2d, src/containers.h:31-43 */
    double y, y_middle, z, z_middle, py, pz, py_next, pz_next, n, gamma, px;
} opart;

static void put_slice(double *dst3d, int i, const double *src2d, size_t nn) {
    if (dst3d) memcpy(dst3d + (size_t)i * nn, src2d, sizeof(double) * nn);
}

/* output_step, src/System_3d.cpp:629-666. */
static void output_step(const qpo_grid *g, qpo_fields *out, int i, int base, const double *psi,
                        const double *psi_prev, const double *by, const double *bz,
                        const double *rho, const double *jx, const double *jy, const double *jz) {
    const int ny = g->ny, nz = g->nz;
    const size_t nn = (size_t)ny * nz;
    const int io = i - base; /* slice i sits at io in the output arrays */
    if (!out) return;
    if (out->ex && i < g->nx - 1) {
        double *ex = out->ex + (size_t)io * nn;
        for (size_t t = 0; t < nn; ++t) ex[t] = (psi[t] - psi_prev[t]) / g->dx;
    }
    if (out->ey) {
        double *ey = out->ey + (size_t)io * nn;
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 0; k < nz; ++k)
                ey[(size_t)nz * j + k] = -(psi[(size_t)nz * (j + 1) + k] - psi[(size_t)nz * j + k]) / g->dy +
                                         bz[(size_t)nz * j + k];
        for (int k = 0; k < nz; ++k)
            ey[(size_t)nz * (ny - 1) + k] = -(psi[k] - psi[(size_t)nz * (ny - 1) + k]) / g->dy +
                                            bz[(size_t)nz * (ny - 1) + k];
    }
    if (out->ez) {
        double *ez = out->ez + (size_t)io * nn;
        for (int j = 0; j < ny; ++j) {
            for (int k = 0; k < nz - 1; ++k)
                ez[(size_t)nz * j + k] = -(psi[(size_t)nz * j + k + 1] - psi[(size_t)nz * j + k]) / g->dz -
                                         by[(size_t)nz * j + k];
            ez[(size_t)nz * j + nz - 1] = -(psi[(size_t)nz * j] - psi[(size_t)nz * j + nz - 1]) / g->dz -
                                          by[(size_t)nz * j + nz - 1];
        }
    }
    put_slice(out->psi, io, psi, nn);
    put_slice(out->rho, io, rho, nn);
    put_slice(out->jx, io, jx, nn);
    put_slice(out->jy, io, jy, nn);
    put_slice(out->jz, io, jz, nn);
}

/* Periodic solve of fourier.in plus the caller's division by ny, nz
 * (src/System_3d.cpp:356-363, 459-467, 539-544, 567-572). */
static void solve_and_scale(const plan2d *pl, const qpo_grid *g, double D, double *src,
                            double *spec, double *dst) {
    const int ny = g->ny, nz = g->nz;
    solve_poisson_plan(pl, g->dy, g->dz, D, src, spec);
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) dst[(size_t)nz * j + k] = (src[(size_t)nz * j + k]) / ny / nz;
}

/* The loop over slices nx-1 .. i_end (i_end = nx - n_slices, or 0 when
 * n_slices <= 0).  windowed: the 3D inputs and outputs hold only slices
 * i_end .. nx-1, slice i at index i - i_end; otherwise all nx slices. */
static int wake_run_impl(const qpo_grid *g, size_t np_in, const double *y0, const double *z0,
                         const double *py0, const double *pz0, const double *w,
                         const double *a_sqr, const double *rho_b, const double *jx_b,
                         qpo_fields *out, int *stop_slice, int n_slices, int windowed) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nn = (size_t)ny * nz;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
    const double thr = g->psi_threshold - 1;
    const long np = (long)np_in;
    const int K = g->K;
    const double D = g->D;

    opart *P = (opart *)calloc(np_in ? np_in : 1, sizeof(opart));
    for (long p = 0; p < np; ++p) {
        P[p].y = y0[p];
        P[p].z = z0[p];
        P[p].py = py0 ? py0[p] : 0.0;
        P[p].pz = pz0 ? pz0[p] : 0.0;
        P[p].py_next = 0;
        P[p].pz_next = 0;
        P[p].gamma = 0;
        P[p].n = w[p];
    }
    double *buf = (double *)calloc(16 * nn + 2 * (size_t)ny * (nz / 2 + 1), sizeof(double));
    double *psi = buf, *psi_prev = buf + nn, *psi_mid = buf + 2 * nn, *rho = buf + 3 * nn;
    double *jx = buf + 4 * nn, *jy = buf + 5 * nn, *jz = buf + 6 * nn, *jyn = buf + 7 * nn;
    double *jzn = buf + 8 * nn, *djy = buf + 9 * nn, *djz = buf + 10 * nn, *rho_ion = buf + 11 * nn;
    double *src = buf + 12 * nn, *by = buf + 13 * nn, *bz = buf + 14 * nn, *chi = buf + 15 * nn;
    double *spec = buf + 16 * nn;
    plan2d pl;
    plan2d_init(&pl, ny, nz);

    const double half_dy = 0.5 * dy, half_dz = 0.5 * dz;

    const int i_end = n_slices > 0 ? (nx - n_slices > 0 ? nx - n_slices : 0) : 0;
    const int base = windowed ? i_end : 0;
    const int nheld = nx - base;

    /* ex(n.x-1) = 0 (:265-270); psi, by, bz, rho_ion, jy, jz zeroed (:272-316). */
    if (out && out->ex) memset(out->ex + (size_t)(nx - 1 - base) * nn, 0, sizeof(double) * nn);
    if (out && out->by) memset(out->by, 0, sizeof(double) * nn * nheld);
    if (out && out->bz) memset(out->bz, 0, sizeof(double) * nn * nheld);

    /* rho_ion deposition (:318-323). */
#pragma omp parallel for
    for (long p = 0; p < np; ++p) qpo_deposit(P[p].y, P[p].z, -P[p].n, rho_ion, ny, nz, dy, dz, 0, 0);

    int stop_flag = 0;
    int stopped_at = -1;

    for (int i = nx - 1; i >= i_end; --i) {
        const double *asq = a_sqr + (size_t)(i - base) * nn;
        const double *rb = rho_b ? rho_b + (size_t)(i - base) * nn : NULL;
        const double *jb = jx_b ? jx_b + (size_t)(i - base) * nn : NULL;

        /* S1 (:334-339) */
        memcpy(psi_prev, psi, sizeof(double) * nn);
        /* S2 (:341-353) */
#pragma omp parallel for
        for (long t = 0; t < (long)nn; ++t) src[t] = (jb ? jb[t] : 0.0) - (rb ? rb[t] : 0.0) - rho_ion[t];
#pragma omp parallel for
        for (long p = 0; p < np; ++p) qpo_deposit_raw(P[p].y, P[p].z, -P[p].n, src, ny, nz, dy, dz);
        /* S3 (:355-365) */
        solve_and_scale(&pl, g, 0.0, src, spec, psi);
        qpo_normalize_potential(psi, ny, nz);
        qpo_increase_minimum(psi, (int)nn, thr);
        /* S4 (:367-378) */
#pragma omp parallel for
        for (long p = 0; p < np; ++p) {
            opart *q = &P[p];
            double a = qpo_gather(q->y, q->z, asq, ny, nz, dy, dz, 0, 0);
            double ps = qpo_gather(q->y, q->z, psi, ny, nz, dy, dz, 0, 0);
            q->gamma = 0.5 * (1 + q->py * q->py + q->pz * q->pz + a + (1 + ps) * (1 + ps)) / (1 + ps);
            q->px = 0.5 * (1 + q->py * q->py + q->pz * q->pz + a - (1 + ps) * (1 + ps)) / (1 + ps);
        }
        /* S5 (:380-397) */
        for (size_t t = 0; t < nn; ++t) {
            rho[t] = rb ? rb[t] : 0.0;
            jx[t] = jb ? jb[t] : 0.0;
            chi[t] = 0.0;
        }
#pragma omp parallel for
        for (long p = 0; p < np; ++p) {
            opart *q = &P[p];
            double vx = q->px / q->gamma;
            qpo_deposit(q->y, q->z, q->n * vx / (1 - vx), jx, ny, nz, dy, dz, 0, 0);
            qpo_deposit(q->y, q->z, q->n / (1 - vx), rho, ny, nz, dy, dz, 0, 0);
            qpo_deposit(q->y, q->z, -q->n / (1 - vx) / q->gamma, chi, ny, nz, dy, dz, 0, 0);
        }
        put_slice(out ? out->chi : NULL, i - base, chi, nn);
        /* S6 (:399-407): by(i) = by(i+1) is the carried 2D buffer */
        put_slice(out ? out->by : NULL, i - base, by, nn);
        put_slice(out ? out->bz : NULL, i - base, bz, nn);

        if (i == 0) break; /* :409-411 */

        for (int it = 0; it < K; ++it) {
            /* P1 advance momenta (:415-433) */
#pragma omp parallel for
            for (long p = 0; p < np; ++p) {
                opart *q = &P[p];
                double ps = qpo_gather(q->y, q->z, psi, ny, nz, dy, dz, 0, 0);
                double day = qpo_gather_yder(q->y, q->z, asq, ny, nz, dy, dz);
                double daz = qpo_gather_zder(q->y, q->z, asq, ny, nz, dy, dz);
                double dpy = qpo_gather_yder(q->y, q->z, psi, ny, nz, dy, dz);
                double dpz = qpo_gather_zder(q->y, q->z, psi, ny, nz, dy, dz);
                double byp = qpo_gather(q->y, q->z, by, ny, nz, dy, dz, 0, half_dz);
                double bzp = qpo_gather(q->y, q->z, bz, ny, nz, dy, dz, half_dy, 0);
                q->py_next = q->py - dx * 0.5 * day / (1 + ps);
                q->py_next += dx * q->gamma * dpy / (1 + ps);
                q->py_next -= dx * bzp;
                q->pz_next = q->pz - dx * 0.5 * daz / (1 + ps);
                q->pz_next += dx * q->gamma * dpz / (1 + ps);
                q->pz_next += dx * byp;
                if (g_pstar_f32) {  /* precision study: the increment p* - p rounded to single */
                    q->py_next = q->py + (double)(float)(q->py_next - q->py);
                    q->pz_next = q->pz + (double)(float)(q->pz_next - q->pz);
                }
            }
            /* P2 advance half coordinate (:435-443) */
#pragma omp parallel for
            for (long p = 0; p < np; ++p) {
                opart *q = &P[p];
                double ps = qpo_gather(q->y, q->z, psi, ny, nz, dy, dz, 0, 0);
                q->y_middle = q->y + 0.5 * dx * q->py_next / (1 + ps);
                q->z_middle = q->z + 0.5 * dx * q->pz_next / (1 + ps);
                qpo_normalize_coordinates(&q->y_middle, &q->z_middle, g->ly, g->lz);
            }
            /* P3 psi source at the middle positions and psi_middle (:445-469) */
#pragma omp parallel for
            for (long t = 0; t < (long)nn; ++t) src[t] = (jb ? jb[t] : 0.0) - (rb ? rb[t] : 0.0) - rho_ion[t];
#pragma omp parallel for
            for (long p = 0; p < np; ++p)
                qpo_deposit_raw(P[p].y_middle, P[p].z_middle, -P[p].n, src, ny, nz, dy, dz);
            solve_and_scale(&pl, g, 0.0, src, spec, psi_mid);
            qpo_normalize_potential(psi_mid, ny, nz);
            qpo_increase_minimum(psi_mid, (int)nn, thr);
            /* P4 jy, jz at the middle positions (:471-486) */
            memset(jyn, 0, sizeof(double) * nn);
            memset(jzn, 0, sizeof(double) * nn);
#pragma omp parallel for
            for (long p = 0; p < np; ++p) {
                opart *q = &P[p];
                double ps = qpo_gather(q->y_middle, q->z_middle, psi_mid, ny, nz, dy, dz, 0, 0);
                qpo_deposit(q->y_middle, q->z_middle, q->n * q->py_next / (1 + ps), jyn, ny, nz, dy, dz, half_dy, 0);
                qpo_deposit(q->y_middle, q->z_middle, q->n * q->pz_next / (1 + ps), jzn, ny, nz, dy, dz, 0, half_dz);
            }
            /* P5 d/dxi of the transverse currents (:488-495) */
            for (size_t t = 0; t < nn; ++t) {
                djy[t] = (jy[t] - jyn[t]) / dx;
                djz[t] = (jz[t] - jzn[t]) / dx;
            }
            /* P6 gamma, px at the old positions (:497-508) */
#pragma omp parallel for
            for (long p = 0; p < np; ++p) {
                opart *q = &P[p];
                double a = qpo_gather(q->y, q->z, asq, ny, nz, dy, dz, 0, 0);
                double ps = qpo_gather(q->y, q->z, psi, ny, nz, dy, dz, 0, 0);
                double p2 = 0.25 * (q->py + q->py_next) * (q->py + q->py_next) +
                            0.25 * (q->pz + q->pz_next) * (q->pz + q->pz_next);
                q->gamma = 0.5 * (1 + p2 + a + (1 + ps) * (1 + ps)) / (1 + ps);
                q->px = 0.5 * (1 + p2 + a - (1 + ps) * (1 + ps)) / (1 + ps);
            }
            /* P7 jx, rho, susceptibility (:510-527) */
            for (size_t t = 0; t < nn; ++t) {
                rho[t] = rb ? rb[t] : 0.0;
                jx[t] = jb ? jb[t] : 0.0;
                chi[t] = 0.0;
            }
#pragma omp parallel for
            for (long p = 0; p < np; ++p) {
                opart *q = &P[p];
                double vx = q->px / q->gamma;
                qpo_deposit(q->y, q->z, q->n * vx / (1 - vx), jx, ny, nz, dy, dz, 0, 0);
                qpo_deposit(q->y, q->z, q->n / (1 - vx), rho, ny, nz, dy, dz, 0, 0);
                qpo_deposit(q->y, q->z, -q->n / (1 - vx) / q->gamma, chi, ny, nz, dy, dz, 0, 0);
            }
            put_slice(out ? out->chi : NULL, i - base, chi, nn);
            /* P8 B_y (:529-552) */
            for (int j = 0; j < ny; ++j) {
                for (int k = 0; k < nz - 1; ++k)
                    src[(size_t)nz * j + k] = -D * by[(size_t)nz * j + k] -
                                              (jx[(size_t)nz * j + k + 1] - jx[(size_t)nz * j + k]) / dz +
                                              djz[(size_t)nz * j + k];
                src[(size_t)nz * j + nz - 1] = -D * by[(size_t)nz * j + nz - 1] -
                                               (jx[(size_t)nz * j] - jx[(size_t)nz * j + nz - 1]) / dz +
                                               djz[(size_t)nz * j + nz - 1];
            }
            solve_and_scale(&pl, g, D, src, spec, by);
            for (size_t t = 0; t < nn; ++t) {
                if (fabs(by[t]) > g->b_threshold) stop_flag = 1;
                if (isnan(by[t])) stop_flag = 1;
            }
            /* P9 B_z (:554-580); the last row uses bz(i, 0, k) as written at :563 */
            for (int j = 0; j < ny - 1; ++j)
                for (int k = 0; k < nz; ++k)
                    src[(size_t)nz * j + k] = -D * bz[(size_t)nz * j + k] +
                                              (jx[(size_t)nz * (j + 1) + k] - jx[(size_t)nz * j + k]) / dy -
                                              djy[(size_t)nz * j + k];
            for (int k = 0; k < nz; ++k)
                src[(size_t)nz * (ny - 1) + k] = -D * bz[k] + (jx[k] - jx[(size_t)nz * (ny - 1) + k]) / dy -
                                                 djy[(size_t)nz * (ny - 1) + k];
            solve_and_scale(&pl, g, D, src, spec, bz);
            for (size_t t = 0; t < nn; ++t) {
                if (fabs(bz[t]) > g->b_threshold) stop_flag = 1;
                if (isnan(bz[t])) stop_flag = 1;
            }
            put_slice(out ? out->by : NULL, i - base, by, nn);
            put_slice(out ? out->bz : NULL, i - base, bz, nn);
            if (stop_flag) break; /* :582-584 */
        }
        if (stop_flag) { /* :588-591 */
            stopped_at = i;
            break;
        }
        /* A1 (:593-598) */
        for (long p = 0; p < np; ++p) {
            P[p].py = P[p].py_next;
            P[p].pz = P[p].pz_next;
        }
        /* A2 advance coordinate (:600-608) */
#pragma omp parallel for
        for (long p = 0; p < np; ++p) {
            opart *q = &P[p];
            double ps = qpo_gather(q->y, q->z, psi_mid, ny, nz, dy, dz, 0, 0);
            q->y += dx * q->py / (1 + ps);
            q->z += dx * q->pz / (1 + ps);
            qpo_normalize_coordinates(&q->y, &q->z, g->ly, g->lz);
        }
        /* A3 output_step(i) (:610) */
        output_step(g, out, i, base, psi, psi_prev, by, bz, rho, jx, jy, jz);
        /* A4 swap (:612-618) */
        memcpy(jy, jyn, sizeof(double) * nn);
        memcpy(jz, jzn, sizeof(double) * nn);
    }

    /* output_step(0) after the loop (:622).  After a stop, by(0)/bz(0) are
     * still the zeros set at the start of the pass. */
    if (i_end == 0) {
        if (stop_flag) {
            memset(by, 0, sizeof(double) * nn);
            memset(bz, 0, sizeof(double) * nn);
        }
        output_step(g, out, 0, base, psi, psi_prev, by, bz, rho, jx, jy, jz);
    }

    plan2d_free(&pl);
    free(buf);
    free(P);
    if (stop_slice) *stop_slice = stopped_at;
    return stop_flag ? 1 : 0;
}

int qpo_wake_run_partial(const qpo_grid *g, size_t np, const double *y, const double *z,
                         const double *py, const double *pz, const double *w,
                         const double *a_sqr, const double *rho_b, const double *jx_b,
                         qpo_fields *out, int *stop_slice, int n_slices) {
    return wake_run_impl(g, np, y, z, py, pz, w, a_sqr, rho_b, jx_b, out, stop_slice, n_slices, 0);
}

int qpo_wake_run_window(const qpo_grid *g, size_t np, const double *y, const double *z,
                        const double *py, const double *pz, const double *w,
                        const double *a_sqr, const double *rho_b, const double *jx_b,
                        qpo_fields *out, int *stop_slice, int n_slices) {
    return wake_run_impl(g, np, y, z, py, pz, w, a_sqr, rho_b, jx_b, out, stop_slice, n_slices, 1);
}

int qpo_wake_run(const qpo_grid *g, size_t np, const double *y, const double *z,
                 const double *py, const double *pz, const double *w,
                 const double *a_sqr, const double *rho_b, const double *jx_b,
                 qpo_fields *out, int *stop_slice) {
    return wake_run_impl(g, np, y, z, py, pz, w, a_sqr, rho_b, jx_b, out, stop_slice, 0, 0);
}


/* ------------------------------------------------------------------------
 * Driver bunch: 3D CIC on the nodes, field gather, explicit push.
 * ---------------------------------------------------------------------- */
static int wrap_index(int j, int n) {
    j %= n;
    return j < 0 ? j + n : j;
}

/* src/array_utils.cpp:290-338, into an array holding slices base .. base+nheld-1
 * (slice i at index i - base); contributions to other slices are dropped. */
static void deposit3d_win(const qpo_grid *g, int base, int nheld, double x, double y, double z, double v,
                          double *a3) {
    const int i1 = (int)floor(x / g->dx), i2 = i1 + 1;
    const double x_frac = x / g->dx - i1;
    if (i1 < 0 || i2 >= g->nx) return;
    const int n1 = g->ny, n2 = g->nz;
    int j1 = (int)floor(y / g->dy);
    const double y_frac = y / g->dy - j1;
    const int j2 = wrap_index(j1 + 1, n1);
    j1 = wrap_index(j1, n1);
    int k1 = (int)floor(z / g->dz);
    const double z_frac = z / g->dz - k1;
    const int k2 = wrap_index(k1 + 1, n2);
    k1 = wrap_index(k1, n2);
    const size_t nn = (size_t)n1 * n2;
    if (i1 >= base && i1 < base + nheld) {
        double *a = a3 + (size_t)(i1 - base) * nn;
        a[(size_t)j1 * n2 + k1] += v * (1 - x_frac) * (1 - y_frac) * (1 - z_frac);
        a[(size_t)j2 * n2 + k1] += v * (1 - x_frac) * y_frac * (1 - z_frac);
        a[(size_t)j1 * n2 + k2] += v * (1 - x_frac) * (1 - y_frac) * z_frac;
        a[(size_t)j2 * n2 + k2] += v * (1 - x_frac) * y_frac * z_frac;
    }
    if (i2 >= base && i2 < base + nheld) {
        double *b = a3 + (size_t)(i2 - base) * nn;
        b[(size_t)j1 * n2 + k1] += v * x_frac * (1 - y_frac) * (1 - z_frac);
        b[(size_t)j2 * n2 + k1] += v * x_frac * y_frac * (1 - z_frac);
        b[(size_t)j1 * n2 + k2] += v * x_frac * (1 - y_frac) * z_frac;
        b[(size_t)j2 * n2 + k2] += v * x_frac * y_frac * z_frac;
    }
}

void qpo_deposit3d(const qpo_grid *g, double x, double y, double z, double v, double *a3) {
    deposit3d_win(g, 0, g->nx, x, y, z, v, a3);
}

/* src/array_utils.cpp:340-374 */
double qpo_gather3d(const qpo_grid *g, double ox, double oy, double oz, double x, double y, double z,
                    const double *a3) {
    const int i1 = (int)floor((x - ox) / g->dx), i2 = i1 + 1;
    const double x_frac = (x - ox) / g->dx - i1;
    if (i1 < 0 || i2 >= g->nx) return 0.0;
    const int n1 = g->ny, n2 = g->nz;
    int j1 = (int)floor((y - oy) / g->dy);
    const double y_frac = (y - oy) / g->dy - j1;
    const int j2 = wrap_index(j1 + 1, n1);
    j1 = wrap_index(j1, n1);
    int k1 = (int)floor((z - oz) / g->dz);
    const double z_frac = (z - oz) / g->dz - k1;
    const int k2 = wrap_index(k1 + 1, n2);
    k1 = wrap_index(k1, n2);
    const size_t nn = (size_t)n1 * n2;
    const double *a = a3 + (size_t)i1 * nn, *b = a3 + (size_t)i2 * nn;
    return a[(size_t)j1 * n2 + k1] * (1 - x_frac) * (1 - y_frac) * (1 - z_frac) +
           a[(size_t)j2 * n2 + k1] * (1 - x_frac) * y_frac * (1 - z_frac) +
           a[(size_t)j1 * n2 + k2] * (1 - x_frac) * (1 - y_frac) * z_frac +
           a[(size_t)j2 * n2 + k2] * (1 - x_frac) * y_frac * z_frac +
           b[(size_t)j1 * n2 + k1] * x_frac * (1 - y_frac) * (1 - z_frac) +
           b[(size_t)j2 * n2 + k1] * x_frac * y_frac * (1 - z_frac) +
           b[(size_t)j1 * n2 + k2] * x_frac * (1 - y_frac) * z_frac +
           b[(size_t)j2 * n2 + k2] * x_frac * y_frac * z_frac;
}

/* src/System_3d.cpp:926-946 */
void qpo_bunch_deposit_window(const qpo_grid *g, int base, int nheld, size_t np, const double *x,
                              const double *y, const double *z, const double *px, const double *gamma,
                              const double *q, double *rho_b, double *jx_b) {
    const size_t vol = (size_t)nheld * g->ny * g->nz;
    memset(rho_b, 0, sizeof(double) * vol);
    memset(jx_b, 0, sizeof(double) * vol);
    for (size_t p = 0; p < np; ++p) {
        deposit3d_win(g, base, nheld, x[p], y[p], z[p], q[p], rho_b);
        deposit3d_win(g, base, nheld, x[p], y[p], z[p], q[p] * px[p] / gamma[p], jx_b);
    }
}

void qpo_bunch_deposit(const qpo_grid *g, size_t np, const double *x, const double *y, const double *z,
                       const double *px, const double *gamma, const double *q, double *rho_b, double *jx_b) {
    qpo_bunch_deposit_window(g, 0, g->nx, np, x, y, z, px, gamma, q, rho_b, jx_b);
}

/* src/System_3d.cpp:843-860 */
void qpo_bunch_fields(const qpo_grid *g, size_t np, const double *x, const double *y, const double *z,
                      const double *px, const double *py, const double *pz, const double *gamma,
                      const double *ex3, const double *ey3, const double *ez3, const double *by3,
                      const double *bz3, double field_schwinger, double *ex, double *ey, double *ez,
                      double *by, double *bz, double *chi) {
    for (size_t p = 0; p < np; ++p) {
        const double hx = 0.5 * g->dx, hy = 0.5 * g->dy, hz = 0.5 * g->dz;  /* :179-183 */
        ex[p] = qpo_gather3d(g, -hx, 0, 0, x[p], y[p], z[p], ex3);
        ey[p] = qpo_gather3d(g, 0, hy, 0, x[p], y[p], z[p], ey3);
        ez[p] = qpo_gather3d(g, 0, 0, hz, x[p], y[p], z[p], ez3);
        by[p] = qpo_gather3d(g, 0, 0, hz, x[p], y[p], z[p], by3);
        bz[p] = qpo_gather3d(g, 0, hy, 0, x[p], y[p], z[p], bz3);
        const double fx = gamma[p] * ex[p] + py[p] * bz[p] - pz[p] * by[p];
        const double fy = gamma[p] * ey[p] - px[p] * bz[p];
        const double fz = gamma[p] * ez[p] + px[p] * by[p];
        const double fl = px[p] * ex[p] + py[p] * ey[p] + pz[p] * ez[p];
        chi[p] = sqrt(fx * fx + fy * fy + fz * fz - fl * fl) / field_schwinger;
    }
}

/* src/System_3d.cpp:862-890 (the QED branch :892-921 is not restated) */
void qpo_bunch_push(size_t np, double *x, double *y, double *z, double *px, double *py, double *pz,
                    double *gamma, const double *cmr, const int *photon, const double *ex,
                    const double *ey, const double *ez, const double *by, const double *bz, double dt) {
    for (size_t p = 0; p < np; ++p) {
        if (photon[p]) {
            x[p] += dt * (px[p] / gamma[p] - 1);
            y[p] += dt * py[p] / gamma[p];
            z[p] += dt * pz[p] / gamma[p];
            gamma[p] = sqrt(px[p] * px[p] + py[p] * py[p] + pz[p] * pz[p]);
            continue;
        }
        const double fx = cmr[p] * (ex[p] + (py[p] * bz[p] - pz[p] * by[p]) / gamma[p]);
        const double fy = cmr[p] * (ey[p] - px[p] * bz[p] / gamma[p]);
        const double fz = cmr[p] * (ez[p] + px[p] * by[p] / gamma[p]);
        x[p] += dt * (px[p] / gamma[p] - 1);
        y[p] += dt * py[p] / gamma[p];
        z[p] += dt * pz[p] / gamma[p];
        px[p] += dt * fx;
        py[p] += dt * fy;
        pz[p] += dt * fz;
        gamma[p] = sqrt(1 + px[p] * px[p] + py[p] * py[p] + pz[p] * pz[p]);
    }
}
