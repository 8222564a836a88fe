#include "System_3d.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <iomanip>
#include <random>
#include <sstream>
#include <stdexcept>

#include "Fourier2d.h"

namespace {

void check(int st, const char *what) {
    if (st < 0) throw std::runtime_error(std::string(what) + ": " + qp_last_error());
}

std::string memory_formatter(long long bytes) {
    double r = (double)bytes;
    const char *pre[] = {"  B", "KiB", "MiB", "GiB", "TiB", "PiB"};
    int i = 0;
    while (r > 1024 && i < 5) {
        r /= 1024;
        ++i;
    }
    std::ostringstream s;
    s << std::setw(5) << std::fixed << std::setprecision(1) << r << " " << pre[i];
    return s.str();
}

// Trilinear deposit onto a 3D array with its origin; x outside the box is
// dropped, y and z are periodic (src/array_utils.cpp:290-338).
void deposit3d(double x, double y, double z, double v, array3d &a) {
    const vector3d s = a.get_steps(), o = a.get_origin();
    const ivector3d n = a.get_dimensions();
    const int i1 = (int)std::floor((x - o.x) / s.x), i2 = i1 + 1;
    const double xf = (x - o.x) / s.x - i1;
    if (i1 < 0 || i2 >= n.x) return;
    int j1 = (int)std::floor((y - o.y) / s.y);
    const int j2 = ((j1 + 1) % n.y + n.y) % n.y;
    const double yf = (y - o.y) / s.y - j1;
    j1 = (j1 % n.y + n.y) % n.y;
    int k1 = (int)std::floor((z - o.z) / s.z);
    const int k2 = ((k1 + 1) % n.z + n.z) % n.z;
    const double zf = (z - o.z) / s.z - k1;
    k1 = (k1 % n.z + n.z) % n.z;
    a(i1, j1, k1) += v * (1 - xf) * (1 - yf) * (1 - zf);
    a(i1, j2, k1) += v * (1 - xf) * yf * (1 - zf);
    a(i1, j1, k2) += v * (1 - xf) * (1 - yf) * zf;
    a(i1, j2, k2) += v * (1 - xf) * yf * zf;
    a(i2, j1, k1) += v * xf * (1 - yf) * (1 - zf);
    a(i2, j2, k1) += v * xf * yf * (1 - zf);
    a(i2, j1, k2) += v * xf * (1 - yf) * zf;
    a(i2, j2, k2) += v * xf * yf * zf;
}

// Trilinear gather, 0 outside the box in x (src/array_utils.cpp:340-374).
double gather3d(double x, double y, double z, const array3d &a) {
    const vector3d s = a.get_steps(), o = a.get_origin();
    const ivector3d n = a.get_dimensions();
    const int i1 = (int)std::floor((x - o.x) / s.x), i2 = i1 + 1;
    const double xf = (x - o.x) / s.x - i1;
    if (i1 < 0 || i2 >= n.x) return 0.0;
    int j1 = (int)std::floor((y - o.y) / s.y);
    const int j2 = ((j1 + 1) % n.y + n.y) % n.y;
    const double yf = (y - o.y) / s.y - j1;
    j1 = (j1 % n.y + n.y) % n.y;
    int k1 = (int)std::floor((z - o.z) / s.z);
    const int k2 = ((k1 + 1) % n.z + n.z) % n.z;
    const double zf = (z - o.z) / s.z - k1;
    k1 = (k1 % n.z + n.z) % n.z;
    return a(i1, j1, k1) * (1 - xf) * (1 - yf) * (1 - zf) + a(i1, j2, k1) * (1 - xf) * yf * (1 - zf) +
           a(i1, j1, k2) * (1 - xf) * (1 - yf) * zf + a(i1, j2, k2) * (1 - xf) * yf * zf +
           a(i2, j1, k1) * xf * (1 - yf) * (1 - zf) + a(i2, j2, k1) * xf * yf * (1 - zf) +
           a(i2, j1, k2) * xf * (1 - yf) * zf + a(i2, j2, k2) * xf * yf * zf;
}

// Synchrotron emission probability rate with the Airy-function approximations
// of src/monte_carlo.cpp:11-37 (QED photon emission of the driver bunch).
double emission_rate(double g, double r, double chi, double omega) {
    const double alpha = ELEMENTARY_CHARGE_CGS * ELEMENTARY_CHARGE_CGS / PLANCK_CONST_BAR_CGS / SPEED_OF_LIGHT_CGS;
    if ((1.0 - r) * g < 1.0 || chi == 0.0 || r == 0.0) return 0.0;
    double rm = 1.0;
    if (chi <= 0.13333) {
        rm = chi / 0.13333;
        r *= rm;
    }
    const double coef = ELECTRON_MASS_CGS * SPEED_OF_LIGHT_CGS * SPEED_OF_LIGHT_CGS / (PLANCK_CONST_BAR_CGS * omega);
    const double y = r / ((1.0 - r) * chi);
    const double x = std::pow(y, 2.0 / 3.0);
    const double e = std::exp(-2.0 * y / 3.0);
    const double ai = e / (2.0 * std::sqrt(QP_PI)) * std::pow(x + 0.80049, -0.75);
    const double dai = -e / (2.0 * std::sqrt(QP_PI)) * std::pow(x + 0.70861 * (1.0 - 0.65 * x / (1.0 + x * x)), 0.25);
    return -rm * alpha * coef / (g * g) * (ai + (2.0 / x + chi * r * std::sqrt(x)) * dai);
}

}  // namespace

// src/System_3d.cpp:64-197
System_3d::System_3d(System_parameters &params, std::ostream &out)
    : l(params.l), dt(params.dt), ppcy(params.ppcy), ppcz(params.ppcz), plasma_profile(params.plasma_profile),
      magnetic_field_iterations(params.magnetic_field_iterations), plasma_units(params.base_frequency_SI),
      output_parameters(params.output_parameters), device(params.device), out(out), species(params.species),
      qed(params.qed), qed_rng(params.device.seed + 7) {
    if (magnetic_field_iterations < 0) throw std::invalid_argument("Magnetic field iterations should be non-negative");
    if (params.ppcy <= 0) throw std::invalid_argument("ppcy should be positive");
    if (params.ppcz <= 0) throw std::invalid_argument("ppcz should be positive");
    n.x = (int)(l.x / params.d.x);
    n.y = (int)(l.y / params.d.y);
    n.z = (int)(l.z / params.d.z);
    out << "Spatial dimensions:  [" << n.x << ", " << n.y << ", " << n.z << "]" << std::endl;
    d.x = l.x / n.x;
    d.y = l.y / n.y;
    d.z = l.z / n.z;
    out << "Steps:               [" << d.x << ", " << d.y << ", " << d.z << "]" << std::endl;
    out << "Simulation box size: [" << l.x << ", " << l.y << ", " << l.z << "]" << std::endl;
    time_iterations = (int)(params.t_end / dt) + 1;
    t_end = dt * (time_iterations - 1);
    out << "Timestep: " << dt << ", end time: " << t_end << ", iterations: " << time_iterations << std::endl;
    out << "----------------------------------------" << std::endl << "PLASMA UNITS\n";
    out << "Frequency " << plasma_units.frequency << " rad/s\nDensity " << plasma_units.density
        << " cm^-3\nWavelength " << plasma_units.wavelength << " cm\nSchwinger field in plasma units: "
        << plasma_units.field_schwinger << "\n----------------------------------------" << std::endl;
    out << "REGISTERED SPECIES\n";
    for (const auto &s : species.get_species())
        out << (s.is_photon ? s.name + " (is photon)" : s.name) << ": charge " << s.charge << ", mass " << s.mass
            << ", charge_to_mass_ratio " << s.charge_to_mass_ratio << "\n";
    out << "----------------------------------------" << std::endl;

    std::vector<size_t> bcount;
    size_t btotal = 0;
    for (const auto &b : params.bunch_parameters_array) {
        bcount.push_back(count_bunch_particles(b.ppc, b.rho));
        btotal += bcount.back();
    }
    const size_t wcount = count_wake_particles();
    const long long nn = (long long)n.y * n.z;
    out << "Expected RAM usage:\n"
        << "3D arrays:       " << memory_formatter(9LL * 8 * n.x * nn) << "\n"
        << "2D arrays:       " << memory_formatter(5LL * 8 * nn) << "\n"
        << "Wake particles:  " << memory_formatter((long long)(5 * 8 * wcount)) << "\n"
        << "Bunch particles: " << memory_formatter((long long)(sizeof(bunch_particle_3d) * btotal)) << "\n"
        << "Device (slice loop): " << memory_formatter(26LL * 8 * nn + 64LL * (long long)wcount) << std::endl;
    out << "----------------------------------------" << std::endl;

    bunch_particles.resize(btotal);
    size_t index = 0;
    for (size_t i = 0; i < params.bunch_parameters_array.size(); ++i) {
        init_bunch_particles(index, params.bunch_parameters_array[i]);
        index += bcount[i];
    }

    const ivector2d syz{n.y, n.z};
    a_sqr = array3d(n, d);
    susceptibility = array3d(n, d);
    rho_bunch = array3d(n, d);
    jx_bunch = array3d(n, d);
    by = array3d(n, d, {0, 0, 0.5 * d.z});
    bz = array3d(n, d, {0, 0.5 * d.y, 0});
    ex = array3d(n, d, {-0.5 * d.x, 0, 0});
    ey = array3d(n, d, {0, 0.5 * d.y, 0});
    ez = array3d(n, d, {0, 0, 0.5 * d.z});
    psi = array2d(syz, d, {0, 0, 0}, Plane::YZ);
    rho = array2d(syz, d, {0, 0, 0}, Plane::YZ);
    jx = array2d(syz, d, {0, 0, 0}, Plane::YZ);
    jy = array2d(syz, d, {-0.5 * d.x, 0.5 * d.y, 0}, Plane::YZ);
    jz = array2d(syz, d, {-0.5 * d.x, 0, 0.5 * d.z}, Plane::YZ);
    init_a_sqr(params.a_sqr);

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
    wp.D = magnetic_field_D;
    wp.psi_threshold = psi_threshold;
    wp.b_threshold = 100.0;
    wp.precision = device.precision;
    wp.use_graph = device.use_graph;
    wp.resort_every = device.resort_every;
    wp.async_output = (device.precision == QP_PREC_MIXED && device.stream_output) ? 1 : 0;
    check(qp_wake_create(qp_host_context(), &wp, &wake), "qp_wake_create");
    check(qp_wake_set_a_sqr(wake, a_sqr.data()), "qp_wake_set_a_sqr");
    if (params.laser_evolve) {
        // a = sqrt(2 <a^2>), real: the deck's laser as the initial envelope
        check(qp_envelope_create(qp_host_context(), n.x, n.y, n.z, d.x, d.y, d.z, params.laser_k0, &env),
              "qp_envelope_create");
        std::vector<double> re((size_t)n.x * n.y * n.z);
        for (size_t t = 0; t < re.size(); ++t) re[t] = std::sqrt(2.0 * a_sqr.data()[t]);
        check(qp_envelope_set(env, re.data(), nullptr), "qp_envelope_set");
        check(qp_wake_attach_envelope(wake, env), "qp_wake_attach_envelope");
        env_chi_iterations = params.laser_chi_iterations;
        out << "Laser envelope: evolving, k0 = " << params.laser_k0 << std::endl;
    }
    if (!bunch_particles.empty() && !qed && device.bunch_on_device) {
        check(qp_bunch_create(qp_host_context(), n.x, n.y, n.z, d.x, d.y, d.z, plasma_units.field_schwinger,
                              &dbunch),
              "qp_bunch_create");
        upload_bunch();
        check(qp_wake_attach_bunch(wake, dbunch), "qp_wake_attach_bunch");
        out << "Driver bunch: on the device (streamed per slice)" << std::endl;
    }
}

// The bunch in qp_bunch's SoA form: q = species charge * n, the species'
// charge-to-mass ratio and photon flag per particle.
void System_3d::upload_bunch() {
    const size_t np = bunch_particles.size();
    std::vector<double> x(np), y(np), z(np), px(np), py(np), pz(np), g(np), q(np), cmr(np);
    std::vector<int> ph(np);
    for (size_t i = 0; i < np; ++i) {
        const bunch_particle_3d &p = bunch_particles[i];
        const auto &sp = species[p.species_id];
        x[i] = p.x;
        y[i] = p.y;
        z[i] = p.z;
        px[i] = p.px;
        py[i] = p.py;
        pz[i] = p.pz;
        g[i] = p.gamma;
        q[i] = sp.charge * p.n;
        cmr[i] = sp.charge_to_mass_ratio;
        ph[i] = sp.is_photon ? 1 : 0;
    }
    check(qp_bunch_set_particles(dbunch, np, x.data(), y.data(), z.data(), px.data(), py.data(), pz.data(),
                                 g.data(), q.data(), cmr.data(), ph.data()),
          "qp_bunch_set_particles");
}

// State and gathered fields back into bunch_particles (upload order), for the writer.
void System_3d::download_bunch() {
    const size_t np = bunch_particles.size();
    std::vector<double> v[13];
    for (auto &a : v) a.resize(np);
    check(qp_bunch_get_particles(dbunch, np, v[0].data(), v[1].data(), v[2].data(), v[3].data(), v[4].data(),
                                 v[5].data(), v[6].data()),
          "qp_bunch_get_particles");
    check(qp_bunch_get_fields(dbunch, np, v[7].data(), v[8].data(), v[9].data(), v[10].data(), v[11].data(),
                              v[12].data()),
          "qp_bunch_get_fields");
    for (size_t i = 0; i < np; ++i) {
        bunch_particle_3d &p = bunch_particles[i];
        p.x = v[0][i];
        p.y = v[1][i];
        p.z = v[2][i];
        p.px = v[3][i];
        p.py = v[4][i];
        p.pz = v[5][i];
        p.gamma = v[6][i];
        p.ex = v[7][i];
        p.ey = v[8][i];
        p.ez = v[9][i];
        p.by = v[10][i];
        p.bz = v[11][i];
        p.chi = v[12][i];
    }
}

System_3d::~System_3d() {
    if (wake) qp_wake_destroy(wake);
    if (dbunch) qp_bunch_destroy(dbunch);
    if (env) qp_envelope_destroy(env);
}

const array3d &System_3d::field3d(const std::string &name) const {
    if (name == "aSqr") return a_sqr;
    if (name == "susceptibility") return susceptibility;
    if (name == "rho_bunch") return rho_bunch;
    if (name == "jx_bunch") return jx_bunch;
    if (name == "ex") return ex;
    if (name == "ey") return ey;
    if (name == "ez") return ez;
    if (name == "by") return by;
    if (name == "bz") return bz;
    throw std::invalid_argument("no 3D array " + name);
}

// src/System_3d.cpp:199-236
void System_3d::run() {
    for (int ti = 0; ti < time_iterations; ++ti) {
        out << "Iteration " << ti << std::endl << "Depositing bunch densities..." << std::endl;
        if (!dbunch) {
            deposit_bunch_densities();
        } else if (output_parameters.output3d) {  // the device streams its own sources per slice
            check(qp_bunch_deposit(dbunch, rho_bunch.data(), jx_bunch.data()), "qp_bunch_deposit");
        }
        if (env) check(qp_envelope_get_a_sqr(env, a_sqr.data()), "qp_envelope_get_a_sqr");  // aSqr output
        out << "Solving wakefield" << std::flush;
        Output_writer writer(output_parameters, ti);
        solve_wakefield(ti, writer);
        out << "Depositing fields to particles..." << std::endl;
        if (dbunch) {
            download_bunch();  // the fields were gathered during the pass
        } else {
            deposit_wakefield_to_bunch();
        }
        writer.write_bunch_parameters(bunch_particles, d.x * d.y * d.z * plasma_units.number_density_norm);
        if (env && ti < time_iterations - 1) {
            out << "Advancing the laser envelope..." << std::endl;
            check(qp_envelope_advance(env, dt, 1, env_chi_iterations, 1, nullptr), "qp_envelope_advance");
        }
        out << "Advancing particles..." << std::endl;
        if (dbunch) {
            if (ti < time_iterations - 1) check(qp_bunch_push(dbunch, dt), "qp_bunch_push");
        } else if (ti < time_iterations - 1) {
            std::vector<bunch_particle_3d> photons;
            double maxp = 0.0;
            advance_bunch_particles(photons, maxp);
            if (qed) {
                out << "Max radiation probability " << maxp << std::endl;
                writer.write_photon_parameters(photons, d.x * d.y * d.z * plasma_units.number_density_norm);
            }
        }
        out << std::endl;
    }
}

int System_3d::slice_callback(const qp_slice_view *v, void *user) {
    auto *self = static_cast<System_3d *>(user);
    const size_t nn = (size_t)self->n.y * self->n.z;
    const size_t b = nn * sizeof(double);
    const int i = v->slice;
    std::memcpy(self->by.slice(i), v->by, b);
    std::memcpy(self->bz.slice(i), v->bz, b);
    std::memcpy(self->susceptibility.slice(i), v->chi, b);
    if (v->kind == QP_SLICE_STOPPED) {
        self->out << "Stop condition encountered, stopping the simulation" << std::endl;
        self->stop_slice_ = i;
        return 0;
    }
    if (i % 10 == 0) self->out << "." << std::flush;
    std::memcpy(self->ex.slice(i), v->ex, b);
    std::memcpy(self->ey.slice(i), v->ey, b);
    std::memcpy(self->ez.slice(i), v->ez, b);
    std::memcpy(self->psi.data(), v->psi, b);
    std::memcpy(self->rho.data(), v->rho, b);
    std::memcpy(self->jx.data(), v->jx, b);
    std::memcpy(self->jy.data(), v->jy, b);
    std::memcpy(self->jz.data(), v->jz, b);
    Output_writer &w = *self->cur_writer;
    w.write_slice(self->psi, "psi", i);
    w.write_slice(self->rho, "rho", i);
    w.write_slice(self->jx, "jx", i);
    w.write_slice(self->jy, "jy", i);
    w.write_slice(self->jz, "jz", i);
    return 0;
}

// src/System_3d.cpp:238-627 on the device
void System_3d::solve_wakefield(int /*iteration*/, Output_writer &writer) {
    const struct {
        const char *name;
        array2d *a;
    } out2d[] = {{"psi", &psi}, {"rho", &rho}, {"jx", &jx}, {"jy", &jy}, {"jz", &jz}};
    for (const auto &o : out2d) writer.initialize_slice_array(n, d, *o.a, o.name);
    init_wake_particles();
    check(qp_wake_set_particles(wake, wy.size(), wy.data(), wz.data(), wpy.data(), wpz.data(), wn.data()),
          "qp_wake_set_particles");
    if (bunch_particles.empty() || dbunch)  // no [bunch], or the attached device bunch streams them
        check(qp_wake_set_bunch_sources(wake, nullptr, nullptr), "qp_wake_set_bunch_sources");
    else
        check(qp_wake_set_bunch_sources(wake, rho_bunch.data(), jx_bunch.data()), "qp_wake_set_bunch_sources");
    cur_writer = &writer;
    stop_slice_ = -1;
    int stop = -1;
    const int st = qp_wake_run(wake, &System_3d::slice_callback, this, &stop);
    check(st, "qp_wake_run");
    cur_writer = nullptr;
    double ms = 0;
    qp_wake_last_time_ms(wake, &ms);
    last_wake_ms_ = ms;
    out << std::endl;
    const struct {
        const char *name;
        array3d *a;
    } out3d[] = {{"aSqr", &a_sqr}, {"susceptibility", &susceptibility}, {"rho_bunch", &rho_bunch},
                 {"jx_bunch", &jx_bunch}, {"ex", &ex}, {"ey", &ey}, {"ez", &ez}, {"by", &by}, {"bz", &bz}};
    for (const auto &o : out3d) writer.write_array(*o.a, o.name);
}

// src/System_3d.cpp:708-722
size_t System_3d::count_wake_particles() const {
    size_t c = 0;
    for (int i = 0; i < ppcy * n.y; ++i)
        for (int j = 0; j < ppcz * n.z; ++j)
            if (plasma_profile((i + 0.5) * d.y / ppcy, (j + 0.5) * d.z / ppcz) != 0) ++c;
    return c;
}

// src/System_3d.cpp:724-748; thermal momenta are an extension ([device] thermal_momentum).
void System_3d::init_wake_particles() {
    wy.clear();
    wz.clear();
    wpy.clear();
    wpz.clear();
    wn.clear();
    const double dppcy = ppcy, dppcz = ppcz;
    std::mt19937_64 rng(device.seed);
    std::normal_distribution<double> th(0.0, device.thermal_momentum > 0 ? device.thermal_momentum : 1.0);
    for (int i = 0; i < ppcy * n.y; ++i)
        for (int j = 0; j < ppcz * n.z; ++j) {
            const double y = (i + 0.5) * d.y / ppcy, z = (j + 0.5) * d.z / ppcz;
            const double v = plasma_profile(y, z);
            if (v == 0) continue;
            wy.push_back(y);
            wz.push_back(z);
            wn.push_back(v / dppcy / dppcz);
            wpy.push_back(device.thermal_momentum > 0 ? th(rng) : 0.0);
            wpz.push_back(device.thermal_momentum > 0 ? th(rng) : 0.0);
        }
}

// src/System_3d.cpp:750-769
size_t System_3d::count_bunch_particles(ivector3d ppc, const profile3d &rhof) const {
    size_t c = 0;
    for (int i = 0; i < ppc.x * n.x; ++i)
        for (int j = 0; j < ppc.y * n.y; ++j)
            for (int k = 0; k < ppc.z * n.z; ++k)
                if (rhof((i + 0.5) * d.x / ppc.x, (j + 0.5) * d.y / ppc.y, (k + 0.5) * d.z / ppc.z) > 0.0) ++c;
    return c;
}

// src/System_3d.cpp:771-803
void System_3d::init_bunch_particles(size_t index, const Bunch_parameters &b) {
    for (int i = 0; i < b.ppc.x * n.x; ++i)
        for (int j = 0; j < b.ppc.y * n.y; ++j)
            for (int k = 0; k < b.ppc.z * n.z; ++k) {
                const double x = (i + 0.5) * d.x / b.ppc.x, y = (j + 0.5) * d.y / b.ppc.y,
                             z = (k + 0.5) * d.z / b.ppc.z;
                const double v = b.rho(x, y, z);
                if (!(v > 0.0)) continue;
                bunch_particle_3d &p = bunch_particles[index++];
                p.x = x;
                p.y = y;
                p.z = z;
                p.py = p.pz = 0;
                p.species_id = b.species_id;
                p.n = v / b.ppc.x / b.ppc.y / b.ppc.z;
                p.gamma = b.gamma;
                p.px = species[b.species_id].is_photon ? b.gamma : std::sqrt(b.gamma * b.gamma - 1);
            }
}

// src/System_3d.cpp:805-816
void System_3d::init_a_sqr(const profile3d &f) {
    for (int i = 0; i < n.x; ++i)
        for (int j = 0; j < n.y; ++j)
            for (int k = 0; k < n.z; ++k) a_sqr(i, j, k) = f(i * d.x, j * d.y, k * d.z);
}

// src/System_3d.cpp:843-860
void System_3d::deposit_wakefield_to_bunch() {
    for (auto &p : bunch_particles) {
        p.ex = gather3d(p.x, p.y, p.z, ex);
        p.ey = gather3d(p.x, p.y, p.z, ey);
        p.ez = gather3d(p.x, p.y, p.z, ez);
        p.by = gather3d(p.x, p.y, p.z, by);
        p.bz = gather3d(p.x, p.y, p.z, bz);
        const double fx = p.gamma * p.ex + p.py * p.bz - p.pz * p.by;
        const double fy = p.gamma * p.ey - p.px * p.bz;
        const double fz = p.gamma * p.ez + p.px * p.by;
        const double fl = p.px * p.ex + p.py * p.ey + p.pz * p.ez;
        p.chi = std::sqrt(fx * fx + fy * fy + fz * fz - fl * fl) / plasma_units.field_schwinger;
    }
}

// src/System_3d.cpp:862-924.  QED emission draws from a seeded generator
// (the reference calls the unseeded rand()), and the recoil subtracts each
// photon momentum component from the same component.
void System_3d::advance_bunch_particles(std::vector<bunch_particle_3d> &photons, double &maxp) {
    std::mt19937_64 &rng = qed_rng;
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (auto &p : bunch_particles) {
        if (species[p.species_id].is_photon) {
            p.x += dt * (p.px / p.gamma - 1);
            p.y += dt * p.py / p.gamma;
            p.z += dt * p.pz / p.gamma;
            p.gamma = std::sqrt(p.px * p.px + p.py * p.py + p.pz * p.pz);
            continue;
        }
        const double cmr = species[p.species_id].charge_to_mass_ratio;
        const double fx = cmr * (p.ex + (p.py * p.bz - p.pz * p.by) / p.gamma);
        const double fy = cmr * (p.ey - p.px * p.bz / p.gamma);
        const double fz = cmr * (p.ez + p.px * p.by / p.gamma);
        p.x += dt * (p.px / p.gamma - 1);
        p.y += dt * p.py / p.gamma;
        p.z += dt * p.pz / p.gamma;
        p.px += dt * fx;
        p.py += dt * fy;
        p.pz += dt * fz;
        p.gamma = std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
        if (!qed) continue;
        double r = U(rng);
        const double prob = emission_rate(p.gamma, r, p.chi, plasma_units.frequency) * dt * p.gamma;
        if (prob > maxp) maxp = prob;
        if (U(rng) < prob) {
            if (p.chi < 0.13333) r *= p.chi / 0.13333;
            bunch_particle_3d ph{};
            ph.px = r * p.px;
            ph.py = r * p.py;
            ph.pz = r * p.pz;
            ph.x = p.x;
            ph.y = p.y;
            ph.z = p.z;
            ph.n = p.n;
            ph.gamma = std::sqrt(ph.px * ph.px + ph.py * ph.py + ph.pz * ph.pz);
            photons.push_back(ph);
            p.px -= ph.px;
            p.py -= ph.py;
            p.pz -= ph.pz;
            p.gamma = std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
        }
    }
}

// src/System_3d.cpp:926-946
void System_3d::deposit_bunch_densities() {
    std::fill(rho_bunch.data(), rho_bunch.data() + (size_t)n.x * n.y * n.z, 0.0);
    std::fill(jx_bunch.data(), jx_bunch.data() + (size_t)n.x * n.y * n.z, 0.0);
    for (const auto &p : bunch_particles) {
        const double q = species[p.species_id].charge * p.n;
        deposit3d(p.x, p.y, p.z, q, rho_bunch);
        deposit3d(p.x, p.y, p.z, q * p.px / p.gamma, jx_bunch);
    }
}
