#include "Wake_batch.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>

#include "Config_reader.h"
#include "Fourier2d.h"

namespace {
void check(int st, const char *what) {
    if (st < 0) throw std::runtime_error(std::string(what) + ": " + qp_last_error());
}
}  // namespace

// Grid arithmetic as System_3d's constructor (src/System_3d.cpp:87-110).
Wake_batch::Wake_batch(System_parameters &params, std::ostream &out)
    : l(params.l), dt(params.dt), magnetic_field_iterations(params.magnetic_field_iterations),
      output_parameters(params.output_parameters), device(params.device), out(out) {
    if (params.wakes.empty()) throw std::invalid_argument("Wake_batch needs at least one [[wake]] member");
    if (magnetic_field_iterations < 0) throw std::invalid_argument("Magnetic field iterations should be non-negative");
    if (!params.bunch_parameters_array.empty()) throw std::invalid_argument("[[wake]] batches take no [bunch]");
    n.x = (int)(l.x / params.d.x);
    n.y = (int)(l.y / params.d.y);
    n.z = (int)(l.z / params.d.z);
    d.x = l.x / n.x;
    d.y = l.y / n.y;
    d.z = l.z / n.z;
    time_iterations = (int)(params.t_end / dt) + 1;
    t_end = dt * (time_iterations - 1);
    out << "Spatial dimensions:  [" << n.x << ", " << n.y << ", " << n.z << "]" << std::endl;
    out << "Steps:               [" << d.x << ", " << d.y << ", " << d.z << "]" << std::endl;
    out << "Timestep: " << dt << ", end time: " << t_end << ", iterations: " << time_iterations << std::endl;
    out << "Batch of " << params.wakes.size() << " wakes" << std::endl;
    keep_fields = output_parameters.output3d || output_parameters.output_xy;

    qp_wake_params wp{};
    wp.nx = n.x;
    wp.ny = n.y;
    wp.nz = n.z;
    wp.dx = d.x;
    wp.dy = d.y;
    wp.dz = d.z;
    wp.ly = l.y;
    wp.lz = l.z;
    wp.K = magnetic_field_iterations;
    wp.D = 10.0;
    wp.psi_threshold = 1e-4;
    wp.b_threshold = 100.0;
    wp.precision = QP_PREC_MIXED;
    wp.use_graph = device.use_graph;
    wp.resort_every = device.resort_every;
    wp.async_output = (keep_fields && device.stream_output) ? 1 : 0;
    check(qp_wake_create_batch(qp_host_context(), &wp, (int)params.wakes.size(), &batch), "qp_wake_create_batch");
    members.resize(params.wakes.size());
    for (size_t m = 0; m < members.size(); ++m) {
        members[m].p = params.wakes[m];
        init_member((int)m);
    }
}

Wake_batch::~Wake_batch() {
    if (batch) qp_wake_destroy(batch);
}

// Lattice of init_wake_particles (src/System_3d.cpp:724-748) with the
// member's profile and ppc, thermal momenta as System_3d ([device]
// thermal_momentum), and the member's Gaussian laser evaluated on the device.
void Wake_batch::init_member(int index) {
    Member &m = members[index];
    std::vector<double> y, z, py, pz, w;
    std::mt19937_64 rng(device.seed);
    std::normal_distribution<double> th(0.0, device.thermal_momentum > 0 ? device.thermal_momentum : 1.0);
    for (int i = 0; i < m.p.ppcy * n.y; ++i)
        for (int j = 0; j < m.p.ppcz * n.z; ++j) {
            const double yy = (i + 0.5) * d.y / m.p.ppcy, zz = (j + 0.5) * d.z / m.p.ppcz;
            const double v = m.p.plasma_profile(yy, zz);
            if (v == 0) continue;
            y.push_back(yy);
            z.push_back(zz);
            w.push_back(v / (double)m.p.ppcy / (double)m.p.ppcz);
            py.push_back(device.thermal_momentum > 0 ? th(rng) : 0.0);
            pz.push_back(device.thermal_momentum > 0 ? th(rng) : 0.0);
        }
    m.particles = y.size();
    check(qp_wake_set_particles_at(batch, index, y.size(), y.data(), z.data(), py.data(), pz.data(), w.data()),
          "qp_wake_set_particles_at");
    check(qp_wake_set_laser_gaussian_at(batch, index, m.p.a0, m.p.sigma[0], m.p.sigma[1], m.p.sigma[2], m.p.r0[0],
                                        m.p.r0[1], m.p.r0[2]),
          "qp_wake_set_laser_gaussian_at");
    out << "Member " << index << " (" << m.p.name << "): a0 = " << m.p.a0 << ", " << m.particles << " particles"
        << std::endl;
    if (!keep_fields) return;
    const ivector2d syz{n.y, n.z};
    m.a_sqr = array3d(n, d);
    const profile3d f = laser_a_sqr(m.p.a0, m.p.sigma, m.p.r0);
    for (int i = 0; i < n.x; ++i)
        for (int j = 0; j < n.y; ++j)
            for (int k = 0; k < n.z; ++k) m.a_sqr(i, j, k) = f(i * d.x, j * d.y, k * d.z);
    m.susceptibility = array3d(n, d);
    m.rho_bunch = array3d(n, d);
    m.jx_bunch = array3d(n, d);
    m.by = array3d(n, d, {0, 0, 0.5 * d.z});
    m.bz = array3d(n, d, {0, 0.5 * d.y, 0});
    m.ex = array3d(n, d, {-0.5 * d.x, 0, 0});
    m.ey = array3d(n, d, {0, 0.5 * d.y, 0});
    m.ez = array3d(n, d, {0, 0, 0.5 * d.z});
    m.psi = array2d(syz, d, {0, 0, 0}, Plane::YZ);
    m.rho = array2d(syz, d, {0, 0, 0}, Plane::YZ);
    m.jx = array2d(syz, d, {0, 0, 0}, Plane::YZ);
    m.jy = array2d(syz, d, {-0.5 * d.x, 0.5 * d.y, 0}, Plane::YZ);
    m.jz = array2d(syz, d, {-0.5 * d.x, 0, 0.5 * d.z}, Plane::YZ);
}

// Per member, what System_3d::slice_callback does for a single wake.
int Wake_batch::slice_callback(int member, const qp_slice_view *v, void *user) {
    auto *self = static_cast<Wake_batch *>(user);
    Member &m = self->members[member];
    const size_t b = (size_t)self->n.y * self->n.z * sizeof(double);
    const int i = v->slice;
    std::memcpy(m.by.slice(i), v->by, b);
    std::memcpy(m.bz.slice(i), v->bz, b);
    std::memcpy(m.susceptibility.slice(i), v->chi, b);
    if (v->kind == QP_SLICE_STOPPED) return 0;
    std::memcpy(m.ex.slice(i), v->ex, b);
    std::memcpy(m.ey.slice(i), v->ey, b);
    std::memcpy(m.ez.slice(i), v->ez, b);
    std::memcpy(m.psi.data(), v->psi, b);
    std::memcpy(m.rho.data(), v->rho, b);
    std::memcpy(m.jx.data(), v->jx, b);
    std::memcpy(m.jy.data(), v->jy, b);
    std::memcpy(m.jz.data(), v->jz, b);
    Output_writer &w = *m.writer;
    w.write_slice(m.psi, "psi", i);
    w.write_slice(m.rho, "rho", i);
    w.write_slice(m.jx, "jx", i);
    w.write_slice(m.jy, "jy", i);
    w.write_slice(m.jz, "jz", i);
    return 0;
}

// The time loop of src/System_3d.cpp:199-236 without a bunch: one batch pass
// per step, every member's files numbered by the step.
void Wake_batch::run() {
    std::vector<int> stops(members.size(), -1);
    for (int ti = 0; ti < time_iterations; ++ti) {
        out << "Iteration " << ti << std::endl << "Solving wakefields" << std::endl;
        for (auto &m : members) {
            m.writer.reset(new Output_writer(output_parameters, ti, m.p.name + "_"));
            if (!keep_fields) continue;
            const struct {
                const char *name;
                array2d *a;
            } out2d[] = {{"psi", &m.psi}, {"rho", &m.rho}, {"jx", &m.jx}, {"jy", &m.jy}, {"jz", &m.jz}};
            for (const auto &o : out2d) m.writer->initialize_slice_array(n, d, *o.a, o.name);
        }
        const int st = qp_wake_run_batch(batch, keep_fields ? &Wake_batch::slice_callback : nullptr, this,
                                         stops.data());
        check(st, "qp_wake_run_batch");
        qp_wake_last_time_ms(batch, &last_wake_ms_);
        for (size_t k = 0; k < members.size(); ++k) {
            Member &m = members[k];
            m.stop = stops[k];
            if (m.stop >= 0) out << "Member " << k << " (" << m.p.name << "): stop condition at slice " << m.stop << std::endl;
            if (keep_fields) {
                const struct {
                    const char *name;
                    array3d *a;
                } out3d[] = {{"aSqr", &m.a_sqr}, {"susceptibility", &m.susceptibility}, {"rho_bunch", &m.rho_bunch},
                             {"jx_bunch", &m.jx_bunch}, {"ex", &m.ex}, {"ey", &m.ey}, {"ez", &m.ez},
                             {"by", &m.by}, {"bz", &m.bz}};
                for (const auto &o : out3d) m.writer->write_array(*o.a, o.name);
            }
            m.writer->write_bunch_parameters({}, 0.0);
            m.writer.reset();
        }
        out << "Wake pass: " << last_wake_ms_ << " ms on the device" << std::endl << std::endl;
    }
}
