// qp_capi.cpp -- the extern "C" boundary (include/quetzal_hip.h).
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "qp_internal.h"
#include "qp_wake.h"

namespace qp {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

}  // namespace qp

using namespace qp;

struct qp_fourier2d_s {
    qp_ctx_s *ctx = nullptr;
    Solver2d S;
    double *d_buf = nullptr;  // n1*n2 real
    double hd1 = 0, hd2 = 0;  // steps the lambda tables were built for
};

#define ARG_CHECK(cond, msg)         \
    do {                             \
        if (!(cond)) {               \
            set_error(msg);          \
            return QP_ERR_ARG;       \
        }                            \
    } while (0)

extern "C" {

const char *qp_last_error(void) { return g_err.c_str(); }
const char *qp_version(void) { return "quetzalpic-amd 0.1 (gfx950)"; }

int qp_ctx_create(int device, qp_ctx *out) {
    ARG_CHECK(out, "qp_ctx_create: null out");
    int n = 0;
    QP_HIP_CHECK(hipGetDeviceCount(&n));
    ARG_CHECK(device >= 0 && device < n, "qp_ctx_create: no such HIP device");
    QP_HIP_CHECK(hipSetDevice(device));
    // once per device, with it current: the address of the allocation-free
    // solve's scratch array on this GPU
    QP_TRY(noalloc_setup(device));
    auto *c = new (std::nothrow) qp_ctx_s();
    if (!c) return QP_ERR_ALLOC;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return QP_ERR_HIP;
    }
    *out = c;
    return QP_OK;
}

int qp_ctx_destroy(qp_ctx ctx) {
    if (!ctx) return QP_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return QP_OK;
}

int qp_ctx_synchronize(qp_ctx ctx) {
    ARG_CHECK(ctx, "null ctx");
    QP_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return QP_OK;
}

// ------------------------------------------------------------ Fourier2d
int qp_fourier2d_create(qp_ctx ctx, int n1, int n2, qp_fourier2d *out) {
    ARG_CHECK(ctx && out, "qp_fourier2d_create: null argument");
    ARG_CHECK(n1 >= 1 && n2 >= 1 && n1 <= kMaxN && n2 <= kMaxN, "qp_fourier2d_create: size out of range [1, 1024]");
    QP_HIP_CHECK(hipSetDevice(ctx->device));
    auto *f = new (std::nothrow) qp_fourier2d_s();
    if (!f) return QP_ERR_ALLOC;
    f->ctx = ctx;
    int st = f->S.init(n1, n2, 1.0, 1.0);
    if (st == QP_OK && hipMalloc(&f->d_buf, sizeof(double) * (size_t)n1 * n2) != hipSuccess) {
        set_error("hipMalloc failed");
        st = QP_ERR_HIP;
    }
    if (st != QP_OK) {
        f->S.destroy();
        delete f;
        return st;
    }
    f->hd1 = f->hd2 = 1.0;
    *out = f;
    return QP_OK;
}

int qp_fourier2d_destroy(qp_fourier2d f) {
    if (!f) return QP_OK;
    (void)hipStreamSynchronize(f->ctx->stream);
    f->S.destroy();
    if (f->d_buf) (void)hipFree(f->d_buf);
    delete f;
    return QP_OK;
}

int qp_fourier2d_forward(qp_fourier2d f, const double *in_real, double *out_complex) {
    ARG_CHECK(f && in_real && out_complex, "qp_fourier2d_forward: null argument");
    hipStream_t s = f->ctx->stream;
    const Solver2d &S = f->S;
    QP_HIP_CHECK(hipMemcpyAsync(f->d_buf, in_real, sizeof(double) * S.n1 * S.n2, hipMemcpyHostToDevice, s));
    QP_TRY(fourier_rows_fwd(S, f->d_buf, S.spec[0], s));
    QP_TRY(solve_cols_forward_only(S, s));
    QP_HIP_CHECK(hipMemcpyAsync(out_complex, S.spec[0], sizeof(double2) * S.n1 * S.nh, hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

int qp_fourier2d_backward(qp_fourier2d f, const double *in_complex, double *out_real) {
    ARG_CHECK(f && in_complex && out_real, "qp_fourier2d_backward: null argument");
    hipStream_t s = f->ctx->stream;
    const Solver2d &S = f->S;
    QP_HIP_CHECK(hipMemcpyAsync(S.spec[0], in_complex, sizeof(double2) * S.n1 * S.nh, hipMemcpyHostToDevice, s));
    QP_TRY(solve_cols_inverse_only(S, s));
    QP_TRY(fourier_rows_inv(S, S.spec[0], f->d_buf, s));
    QP_HIP_CHECK(hipMemcpyAsync(out_real, f->d_buf, sizeof(double) * S.n1 * S.n2, hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

static int ensure_steps(qp_fourier2d f, double d1, double d2) {
    if (f->hd1 != d1 || f->hd2 != d2) {
        QP_HIP_CHECK(hipStreamSynchronize(f->ctx->stream));
        QP_TRY(f->S.set_steps(d1, d2));
        f->hd1 = d1;
        f->hd2 = d2;
    }
    return QP_OK;
}

int qp_solve2d_device(qp_fourier2d f, double d1, double d2, double D, const double *src_dev, double *dst_dev) {
    ARG_CHECK(f && src_dev && dst_dev, "qp_solve2d_device: null argument");
    ARG_CHECK(d1 > 0 && d2 > 0, "qp_solve2d_device: steps must be positive");
    QP_TRY(ensure_steps(f, d1, d2));
    hipStream_t s = f->ctx->stream;
    const Solver2d &S = f->S;
    QP_TRY(solve_rows_fwd_plain(S, src_dev, S.spec[0], s));
    QP_TRY(solve_cols(S, D, 1, s, nullptr, 0));
    QP_TRY(solve_rows_inv_unscaled(S, S.spec[0], dst_dev, s));
    return QP_OK;
}

int qp_solve2d(qp_fourier2d f, double d1, double d2, double D, const double *src, double *dst) {
    ARG_CHECK(f && src && dst, "qp_solve2d: null argument");
    hipStream_t s = f->ctx->stream;
    const size_t b = sizeof(double) * f->S.n1 * f->S.n2;
    QP_HIP_CHECK(hipMemcpyAsync(f->d_buf, src, b, hipMemcpyHostToDevice, s));
    QP_TRY(qp_solve2d_device(f, d1, d2, D, f->d_buf, f->d_buf));
    QP_HIP_CHECK(hipMemcpyAsync(dst, f->d_buf, b, hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    return QP_OK;
}

// ----------------------------------------------------------------- wake
int qp_wake_create(qp_ctx ctx, const qp_wake_params *params, qp_wake *out) {
    return qp_wake_create_batch(ctx, params, 1, out);
}

int qp_wake_create_batch(qp_ctx ctx, const qp_wake_params *params, int nwakes, qp_wake *out) {
    ARG_CHECK(ctx && params && out, "qp_wake_create: null argument");
    ARG_CHECK(nwakes >= 1, "qp_wake_create_batch: nwakes must be >= 1");
    const qp_wake_params &p = *params;
    ARG_CHECK(p.nx >= 1 && p.ny >= 2 && p.nz >= 2, "qp_wake_create: need nx >= 1, ny >= 2, nz >= 2");
    ARG_CHECK(p.ny <= kMaxN && p.nz <= kMaxN, "qp_wake_create: ny, nz must be <= 1024");
    ARG_CHECK(p.ny < 65536 && p.nz < 65536, "qp_wake_create: grid too large");
    ARG_CHECK(p.dx > 0 && p.dy > 0 && p.dz > 0 && p.ly > 0 && p.lz > 0, "qp_wake_create: steps must be positive");
    // Magnetic field iterations should be non-negative (src/System_3d.cpp:77-79)
    ARG_CHECK(p.K >= 0, "Magnetic field iterations should be non-negative");
    ARG_CHECK(p.precision == QP_PREC_FP64 || p.precision == QP_PREC_MIXED, "qp_wake_create: unknown precision");
    ARG_CHECK(nwakes == 1 || p.precision == QP_PREC_MIXED, "qp_wake_create_batch: batches need QP_PREC_MIXED");
    ARG_CHECK(nwakes == 1 || p.resort_every == 0, "qp_wake_create_batch: re-binning is for single wakes only");
    ARG_CHECK(p.longitudinal_b == 0 || p.precision == QP_PREC_MIXED,
              "qp_wake_create: longitudinal_b is an extension of the mixed path");
    ARG_CHECK(p.async_output == 0 || p.precision == QP_PREC_MIXED,
              "qp_wake_create: async_output is a mode of the mixed path");
    ARG_CHECK(p.slice_limit >= 0, "qp_wake_create: slice_limit must be >= 0");
    QP_HIP_CHECK(hipSetDevice(ctx->device));
    auto *h = new (std::nothrow) qp_wake_s();
    if (!h) return QP_ERR_ALLOC;
    Wake *w = &h->w;
    w->ctx = ctx;
    w->p = p;
    w->nw = nwakes;
    w->mem.assign(nwakes, WakeMember{});
    int st = wake_alloc_common(w);
    if (st != QP_OK) {
        qp_wake_destroy(h);
        return st;
    }
    *out = h;
    return QP_OK;
}

int qp_wake_destroy(qp_wake h) {
    if (!h) return QP_OK;
    Wake *w = &h->w;
    (void)hipStreamSynchronize(w->stream());
    if (w->env) --w->env->attached;
    if (w->bunch) --w->bunch->attached;
    wake_free_particles(w);
    wake_destroy_mixed(w);
    w->S.destroy();
    if (w->g_base) (void)hipFree(w->g_base);
    if (w->a_sqr3) (void)hipFree(w->a_sqr3);
    if (w->rho_b3) (void)hipFree(w->rho_b3);
    if (w->jx_b3) (void)hipFree(w->jx_b3);
    if (w->ctl) (void)hipFree(w->ctl);
    if (w->d_laser) (void)hipFree(w->d_laser);
    if (w->h_out) (void)hipHostFree(w->h_out);
    if (w->ev0) (void)hipEventDestroy(w->ev0);
    if (w->ev1) (void)hipEventDestroy(w->ev1);
    for (hipEvent_t e : w->prof_ev) (void)hipEventDestroy(e);
    delete h;
    return QP_OK;
}

int qp_wake_set_particles(qp_wake h, size_t np, const double *y, const double *z, const double *py,
                          const double *pz, const double *weight) {
    return qp_wake_set_particles_at(h, 0, np, y, z, py, pz, weight);
}

int qp_wake_set_particles_at(qp_wake h, int member, size_t np, const double *y, const double *z, const double *py,
                             const double *pz, const double *weight) {
    ARG_CHECK(h, "null wake");
    ARG_CHECK(np == 0 || (y && z && weight), "qp_wake_set_particles: y, z, weight required");
    Wake *w = &h->w;
    ARG_CHECK(member >= 0 && member < w->nw, "qp_wake_set_particles_at: member out of range");
    ARG_CHECK(np < (size_t)INT_MAX / 2, "qp_wake_set_particles: too many particles");
    for (size_t i = 0; i < np; ++i) {
        if (!(y[i] >= 0 && y[i] < w->p.ly && z[i] >= 0 && z[i] < w->p.lz)) {
            set_error("qp_wake_set_particles: particle " + std::to_string(i) + " outside [0, ly) x [0, lz)");
            return QP_ERR_ARG;
        }
    }
    WakeMember &m = w->mem[member];
    m.np = np;
    m.h_y.assign(y, y + np);
    m.h_z.assign(z, z + np);
    m.h_w.assign(weight, weight + np);
    m.h_py.assign(np, 0.0);
    m.h_pz.assign(np, 0.0);
    if (py) m.h_py.assign(py, py + np);
    if (pz) m.h_pz.assign(pz, pz + np);
    m.particles_set = true;
    w->particles_dirty = true;
    return QP_OK;
}

static int upload3d(Wake *w, double **dst, const double *src) {
    const size_t n = (size_t)w->p.nx * w->nn();
    if (!src) {
        if (*dst) (void)hipFree(*dst);
        *dst = nullptr;
        return QP_OK;
    }
    if (!*dst) QP_HIP_CHECK(hipMalloc(dst, sizeof(double) * n));
    QP_HIP_CHECK(hipMemcpy(*dst, src, sizeof(double) * n, hipMemcpyHostToDevice));
    return QP_OK;
}

int qp_wake_set_a_sqr(qp_wake h, const double *a_sqr) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    ARG_CHECK(w->nw == 1, "qp_wake_set_a_sqr: 3D arrays are for single wakes (use the Gaussian laser)");
    ++w->inputs_epoch;
    QP_TRY(upload3d(w, &w->a_sqr3, a_sqr));
    w->mem[0].laser_set = false;
    return QP_OK;
}

int qp_wake_set_laser_gaussian(qp_wake h, double a0, double xsigma, double ysigma, double zsigma, double x0,
                               double y0, double z0) {
    return qp_wake_set_laser_gaussian_at(h, 0, a0, xsigma, ysigma, zsigma, x0, y0, z0);
}

int qp_wake_set_laser_gaussian_at(qp_wake h, int member, double a0, double xsigma, double ysigma, double zsigma,
                                  double x0, double y0, double z0) {
    ARG_CHECK(h, "null wake");
    ARG_CHECK(xsigma > 0 && ysigma > 0 && zsigma > 0, "laser widths must be positive");
    Wake *w = &h->w;
    ARG_CHECK(member >= 0 && member < w->nw, "qp_wake_set_laser_gaussian_at: member out of range");
    const double s8 = std::sqrt(8.0);
    w->mem[member].laser = LaserGauss{0.5 * a0 * a0, xsigma / s8, ysigma / s8, zsigma / s8, x0, y0, z0};
    w->mem[member].laser_set = true;
    ++w->inputs_epoch;
    if (w->a_sqr3) (void)hipFree(w->a_sqr3);
    w->a_sqr3 = nullptr;
    return QP_OK;
}

int qp_wake_set_bunch_sources(qp_wake h, const double *rho_b, const double *jx_b) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    ARG_CHECK(w->nw == 1 || (!rho_b && !jx_b), "qp_wake_set_bunch_sources: bunch sources are for single wakes");
    ++w->inputs_epoch;
    QP_TRY(upload3d(w, &w->rho_b3, rho_b));
    QP_TRY(upload3d(w, &w->jx_b3, jx_b));
    return QP_OK;
}

int qp_wake_attach_bunch(qp_wake h, qp_bunch bunch) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    if (bunch) {
        const Bunch &b = bunch->b;
        ARG_CHECK(w->nw == 1, "qp_wake_attach_bunch: bunches are for single wakes");
        ARG_CHECK(b.ctx == w->ctx, "qp_wake_attach_bunch: bunch and wake must share the context");
        auto same = [](double a, double c) { return std::fabs(a - c) <= 1e-12 * std::fabs(c); };
        ARG_CHECK(b.nx == w->p.nx && b.ny == w->p.ny && b.nz == w->p.nz && same(b.dx, w->p.dx) &&
                      same(b.dy, w->p.dy) && same(b.dz, w->p.dz),
                  "qp_wake_attach_bunch: bunch grid differs from the wake's");
    }
    if (w->bunch) --w->bunch->attached;
    w->bunch = bunch ? &bunch->b : nullptr;
    if (w->bunch) {
        ++w->bunch->attached;
        w->bunch_epoch = w->bunch->epoch;
    }
    ++w->inputs_epoch;
    return QP_OK;
}

int qp_wake_set_envelope_step(qp_wake h, double dtau, int chi_iterations) {
    ARG_CHECK(h, "null wake");
    ARG_CHECK(dtau >= 0 && std::isfinite(dtau), "qp_wake_set_envelope_step: dtau must be >= 0");
    ARG_CHECK(chi_iterations >= 0 && chi_iterations <= 16, "qp_wake_set_envelope_step: chi_iterations in [0, 16]");
    Wake *w = &h->w;
    w->env_dtau = dtau;
    w->env_chi_it = chi_iterations;
    return QP_OK;
}

int qp_wake_attach_envelope(qp_wake h, qp_envelope env) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    if (env) {
        const Envelope &e = env->e;
        ARG_CHECK(w->nw == 1, "qp_wake_attach_envelope: envelopes are for single wakes");
        ARG_CHECK(e.ctx == w->ctx, "qp_wake_attach_envelope: envelope and wake must share the context");
        auto same = [](double a, double b) { return std::fabs(a - b) <= 1e-12 * std::fabs(b); };
        ARG_CHECK(e.nx == w->p.nx && e.ny == w->p.ny && e.nz == w->p.nz && same(e.dx, w->p.dx) &&
                      same(e.dy, w->p.dy) && same(e.dz, w->p.dz),
                  "qp_wake_attach_envelope: envelope grid differs from the wake's");
    }
    if (w->env) --w->env->attached;
    w->env = env ? &env->e : nullptr;
    if (w->env) ++w->env->attached;
    ++w->inputs_epoch;
    return QP_OK;
}

namespace {
struct SingleCb {
    qp_slice_cb cb;
    void *user;
};
int single_cb_adapter(int, const qp_slice_view *v, void *user) {
    const SingleCb *a = static_cast<const SingleCb *>(user);
    return a->cb(v, a->user);
}
bool all_particles_set(const Wake *w) {
    for (const WakeMember &m : w->mem)
        if (!m.particles_set) return false;
    return true;
}
}  // namespace

int qp_wake_run(qp_wake h, qp_slice_cb cb, void *user, int *stop_slice) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    ARG_CHECK(w->nw == 1, "qp_wake_run: batch handle (use qp_wake_run_batch)");
    ARG_CHECK(all_particles_set(w), "qp_wake_run: particles not set");
    QP_HIP_CHECK(hipSetDevice(w->ctx->device));
    if (w->p.precision == QP_PREC_FP64) return wake_run_fp64(w, cb, user, stop_slice);
    SingleCb a{cb, user};
    return wake_run_mixed(w, cb ? single_cb_adapter : nullptr, &a, stop_slice);
}

int qp_wake_run_batch(qp_wake h, qp_batch_cb cb, void *user, int *stop_slices) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    ARG_CHECK(all_particles_set(w), "qp_wake_run_batch: particles not set for every member");
    QP_HIP_CHECK(hipSetDevice(w->ctx->device));
    if (w->p.precision == QP_PREC_FP64) {
        ARG_CHECK(!cb, "qp_wake_run_batch: FP64 handles take qp_wake_run");
        return wake_run_fp64(w, nullptr, nullptr, stop_slices);
    }
    return wake_run_mixed(w, cb, user, stop_slices);
}

// ------------------------------------------------------------- groups
int qp_group_create(qp_ctx ctx, int nbatches, const qp_wake *batches, qp_group *out) {
    ARG_CHECK(ctx && batches && out, "qp_group_create: null argument");
    ARG_CHECK(nbatches >= 1 && nbatches <= 16, "qp_group_create: 1 to 16 batches");
    for (int b = 0; b < nbatches; ++b) {
        ARG_CHECK(batches[b], "qp_group_create: null batch");
        ARG_CHECK(batches[b]->w.ctx == ctx, "qp_group_create: every batch must be on the group's context");
        ARG_CHECK(batches[b]->w.p.precision == QP_PREC_MIXED, "qp_group_create: mixed-precision batches only");
        for (int c = 0; c < b; ++c) ARG_CHECK(batches[c] != batches[b], "qp_group_create: a batch appears twice");
    }
    QP_HIP_CHECK(hipSetDevice(ctx->device));
    auto *h = new (std::nothrow) qp_group_s();
    if (!h) return QP_ERR_ALLOC;
    Group &g = h->g;
    g.ctx = ctx;
    for (int b = 0; b < nbatches; ++b) g.b.push_back(&batches[b]->w);
    if (hipEventCreate(&g.ev0) != hipSuccess || hipEventCreate(&g.ev1) != hipSuccess) {
        if (g.ev0) (void)hipEventDestroy(g.ev0);
        if (g.ev1) (void)hipEventDestroy(g.ev1);
        delete h;
        set_error("qp_group_create: hipEventCreate failed");
        return QP_ERR_HIP;
    }
    *out = h;
    return QP_OK;
}

int qp_group_destroy(qp_group h) {
    if (!h) return QP_OK;
    (void)hipStreamSynchronize(h->g.ctx->stream);
    group_destroy_mixed(&h->g);
    (void)hipEventDestroy(h->g.ev0);
    (void)hipEventDestroy(h->g.ev1);
    delete h;
    return QP_OK;
}

int qp_group_run(qp_group h, qp_group_cb cb, void *user, int *stop_slices) {
    ARG_CHECK(h, "null group");
    for (Wake *w : h->g.b) ARG_CHECK(all_particles_set(w), "qp_group_run: particles not set for every member");
    QP_HIP_CHECK(hipSetDevice(h->g.ctx->device));
    return group_run_mixed(&h->g, cb, user, stop_slices);
}

int qp_group_last_time_ms(qp_group h, double *ms) {
    ARG_CHECK(h && ms, "null argument");
    *ms = h->g.last_ms;
    return QP_OK;
}

int qp_wake_window_stats(qp_wake h, int member, long long *overflow_blocks, long long *deposit_blocks) {
    ARG_CHECK(h && overflow_blocks && deposit_blocks, "null argument");
    Wake *w = &h->w;
    ARG_CHECK(w->p.precision == QP_PREC_MIXED, "qp_wake_window_stats: mixed-precision handles only");
    ARG_CHECK(member >= 0 && member < w->nw, "qp_wake_window_stats: member out of range");
    ARG_CHECK((int)w->win_blocks.size() == w->nw, "qp_wake_window_stats: no pass has run");
    *overflow_blocks = w->win_ovf[member];
    *deposit_blocks = w->win_blocks[member];
    return QP_OK;
}

int qp_wake_batch_size(qp_wake h, int *nwakes) {
    ARG_CHECK(h && nwakes, "null argument");
    *nwakes = h->w.nw;
    return QP_OK;
}

int qp_wake_last_time_ms(qp_wake h, double *ms) {
    ARG_CHECK(h && ms, "null argument");
    *ms = h->w.last_ms;
    return QP_OK;
}

int qp_wake_profile_kernel(qp_wake h, const char *kernel_class, int *launches, double *mean_us) {
    ARG_CHECK(h && kernel_class && launches && mean_us, "null argument");
    Wake *w = &h->w;
    ARG_CHECK(all_particles_set(w), "qp_wake_profile_kernel: particles not set");
    w->prof_class = kernel_class;
    w->prof_count = 0;
    int st;
    if (w->p.precision == QP_PREC_FP64) {
        st = wake_run_fp64(w, nullptr, nullptr, nullptr);
    } else {
        st = wake_run_mixed(w, nullptr, nullptr, nullptr);
    }
    w->prof_class = nullptr;
    if (st < 0) return st;
    QP_HIP_CHECK(hipStreamSynchronize(w->stream()));
    double sum = 0;
    for (int i = 0; i < w->prof_count; ++i) {
        float ms = 0;
        QP_HIP_CHECK(hipEventElapsedTime(&ms, w->prof_ev[2 * i], w->prof_ev[2 * i + 1]));
        sum += ms;
    }
    *launches = w->prof_count;
    *mean_us = w->prof_count ? 1000.0 * sum / w->prof_count : 0.0;
    return QP_OK;
}

int qp_wake_set_tile(qp_wake h, int threads_per_block, int window_nodes) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    ARG_CHECK(w->p.precision == QP_PREC_MIXED, "qp_wake_set_tile: mixed-precision handles only");
    if (threads_per_block == 0 && window_nodes == 0) {
        w->tile_tuned = false;   // the autotuner picks again on the next pass
        return QP_OK;
    }
    ARG_CHECK(mixed_tile_valid(threads_per_block, window_nodes),
              "qp_wake_set_tile: tile not instantiated (see qp_wake_describe for the list)");
    w->tile_pt = threads_per_block;
    w->tile_cap = window_nodes;
    w->tile_tuned = true;
    return QP_OK;
}

int qp_group_set_tile(qp_group h, int threads_per_block, int window_nodes) {
    ARG_CHECK(h, "null group");
    Group *g = &h->g;
    if (threads_per_block == 0 && window_nodes == 0) {
        g->tile_tuned = false;
        return QP_OK;
    }
    ARG_CHECK(mixed_tile_valid(threads_per_block, window_nodes),
              "qp_group_set_tile: tile not instantiated (see qp_wake_describe for the list)");
    g->tile_pt = threads_per_block;
    g->tile_cap = window_nodes;
    g->tile_tuned = true;
    return QP_OK;
}

int qp_group_set_fork(qp_group h, int fork_families) {
    ARG_CHECK(h, "null group");
    h->g.fork_families = fork_families != 0;
    return QP_OK;
}

int qp_wake_describe(qp_wake h, char *buf, size_t len) {
    ARG_CHECK(h && buf && len > 0, "null argument");
    const Wake *w = &h->w;
    std::string s = "{\"members\": " + std::to_string(w->nw) +
                    ", \"tile\": [" + std::to_string(w->tile_pt) + ", " + std::to_string(w->tile_cap) + "]" +
                    ", \"tile_tuned\": " + (w->tile_tuned ? "1" : "0") + ", \"tiles\": " + mixed_tiles_json() +
                    ", \"rows_per_block\": " + std::to_string(w->S.rows_per_block) +
                    ", \"cols_per_block\": " + std::to_string(w->S.cols_per_block) +
                    ", \"precision\": " + std::to_string(w->p.precision) +
                    ", \"fft\": \"" + (w->S.wave ? "wave" : "generic") + "\"" +
                    ", \"envelope\": " + (w->env ? "1" : "0") + ", \"bunch\": " + (w->bunch ? "1" : "0") + "}";
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
    return QP_OK;
}

int qp_wake_get_particles(qp_wake h, size_t np, double *y, double *z, double *py, double *pz) {
    return qp_wake_get_particles_at(h, 0, np, y, z, py, pz);
}

int qp_wake_get_particles_at(qp_wake h, int member, size_t np, double *y, double *z, double *py, double *pz) {
    ARG_CHECK(h, "null wake");
    Wake *w = &h->w;
    ARG_CHECK(member >= 0 && member < w->nw, "qp_wake_get_particles_at: member out of range");
    ARG_CHECK(np == w->mem[member].np, "qp_wake_get_particles: count mismatch");
    if (w->p.precision == QP_PREC_FP64) return wake_download_fp64(w, y, z, py, pz);
    return wake_download_mixed(w, member, y, z, py, pz);
}

}  // extern "C"

// ------------------------------------------------------------ helpers
namespace qp {

int wake_alloc_common(Wake *w) {
    const qp_wake_params &p = w->p;
    const size_t nn = w->nn();
    const size_t nw = (size_t)w->nw;
    QP_TRY(w->S.init(p.ny, p.nz, p.dy, p.dz, w->nw));
    const int ngrids = 27;
    QP_HIP_CHECK(hipMalloc(&w->g_base, sizeof(double) * nn * nw * ngrids));
    QP_HIP_CHECK(hipMemset(w->g_base, 0, sizeof(double) * nn * nw * ngrids));
    double **grids[] = {&w->src, &w->psi, &w->psi_prev, &w->psim, &w->rho, &w->jx, &w->jy, &w->jz, &w->jyn,
                        &w->jzn, &w->djy, &w->djz, &w->rho_ion, &w->by, &w->bz, &w->chi, &w->ex, &w->ey,
                        &w->ez, &w->src2, &w->asq2, &w->psi_b[0], &w->psi_b[1], &w->zero2, &w->bx};
    int gi = 0;
    for (double **g : grids) *g = w->g_base + nn * nw * (gi++);
    QP_HIP_CHECK(hipMalloc(&w->ctl, sizeof(Ctl) * nw));
    QP_HIP_CHECK(hipMemset(w->ctl, 0, sizeof(Ctl) * nw));
    QP_HIP_CHECK(hipMalloc(&w->d_laser, sizeof(LaserGauss) * nw));
    QP_HIP_CHECK(hipMemset(w->d_laser, 0, sizeof(LaserGauss) * nw));
    QP_HIP_CHECK(hipHostMalloc(&w->h_out, sizeof(double) * nn * 12, hipHostMallocDefault));
    QP_HIP_CHECK(hipEventCreate(&w->ev0));
    QP_HIP_CHECK(hipEventCreate(&w->ev1));
    if (p.precision == QP_PREC_MIXED) {
        QP_TRY(w->S.autotune(w->stream(), w->src));
        QP_HIP_CHECK(hipStreamSynchronize(w->stream()));
    }
    return QP_OK;
}

void wake_free_particles(Wake *w) {
    PartFP64 &P = w->pf;
    double *arrs[] = {P.y, P.z, P.py, P.pz, P.n, P.gamma, P.px, P.ym, P.zm, P.pyn, P.pzn};
    for (double *a : arrs)
        if (a) (void)hipFree(a);
    w->pf = PartFP64{};
    for (double *&a : w->init64) {
        if (a) (void)hipFree(a);
        a = nullptr;
    }
}

const double *wake_rho_b(Wake *w, int i) {
    const double *b = w->rho_b_src();
    return b ? b + (size_t)i * w->bsrc_stride() : nullptr;
}
const double *wake_jx_b(Wake *w, int i) {
    const double *b = w->jx_b_src();
    return b ? b + (size_t)i * w->bsrc_stride() : nullptr;
}

void prof_begin(Wake *w, const char *cls) {
    if (!w->prof_class || std::strcmp(cls, w->prof_class) != 0) return;
    const size_t need = 2 * (size_t)(w->prof_count + 1);
    while (w->prof_ev.size() < need) {
        hipEvent_t e;
        // no system-scope fence on record: these events only time device work
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
        w->prof_ev.push_back(e);
    }
    (void)hipEventRecord(w->prof_ev[2 * w->prof_count], w->stream());
}

void prof_end(Wake *w, const char *cls) {
    if (!w->prof_class || std::strcmp(cls, w->prof_class) != 0) return;
    if (w->prof_ev.size() < 2 * (size_t)(w->prof_count + 1)) return;
    (void)hipEventRecord(w->prof_ev[2 * w->prof_count + 1], w->stream());
    w->prof_count++;
}

int deliver_slice(Wake *w, qp_slice_cb cb, void *user, int slice, int kind, const double *psi, const double *rho,
                  const double *jx, const double *jy, const double *jz, const double *ex, const double *ey,
                  const double *ez, const double *by, const double *bz, const double *chi, const double *bx) {
    const size_t nn = w->nn();
    hipStream_t s = w->stream();
    const double *src[12] = {psi, rho, jx, jy, jz, ex, ey, ez, by, bz, chi, bx};
    const int nf = bx ? 12 : 11;
    for (int f = 0; f < nf; ++f)
        QP_HIP_CHECK(hipMemcpyAsync(w->h_out + f * nn, src[f] ? src[f] : w->zero2, sizeof(double) * nn,
                                    hipMemcpyDeviceToHost, s));
    QP_HIP_CHECK(hipStreamSynchronize(s));
    qp_slice_view v;
    v.slice = slice;
    v.kind = kind;
    const double *h = w->h_out;
    v.psi = h;
    v.rho = h + nn;
    v.jx = h + 2 * nn;
    v.jy = h + 3 * nn;
    v.jz = h + 4 * nn;
    v.ex = h + 5 * nn;
    v.ey = h + 6 * nn;
    v.ez = h + 7 * nn;
    v.by = h + 8 * nn;
    v.bz = h + 9 * nn;
    v.chi = h + 10 * nn;
    v.bx = bx ? h + 11 * nn : nullptr;
    int r = cb(&v, user);
    if (r != 0) {
        set_error("slice callback returned " + std::to_string(r));
        return QP_ERR_STATE;
    }
    return QP_OK;
}

}  // namespace qp
