#include "Config_reader.h"

#include <cmath>
#include <sstream>

#include "Plasma_units.h"

namespace {
std::string show(double v) {
    std::ostringstream s;
    s << v;
    return s.str();
}
std::string show(int v) { return std::to_string(v); }
std::string show(bool v) { return v ? "true" : "false"; }
std::string show(const std::string &v) { return v; }
}  // namespace

Config_reader::Config_reader(const std::string &filename, std::ostream &out) : out(out) {
    out << "Reading config " << filename << "\n" << std::endl;
    try {
        config = toml_lite::parse_file(filename);
    } catch (const toml_lite::parse_error &e) {
        throw Config_exception(e.what());
    }
    read_all();
}

Config_reader::Config_reader(toml_lite::TablePtr cfg, std::ostream &out) : config(std::move(cfg)), out(out) {
    read_all();
}

Config_reader Config_reader::from_string(const std::string &text, std::ostream &out) {
    toml_lite::TablePtr t;
    try {
        t = toml_lite::parse_string(text);
    } catch (const toml_lite::parse_error &e) {
        throw Config_exception(e.what());
    }
    return Config_reader(t, out);
}

template <class T>
T Config_reader::read_value(const std::string &name, const toml_lite::TablePtr &table) {
    T v;
    if (table && table->get(name, v)) {
        out << name << " = " << show(v) << std::endl;
        return v;
    }
    throw Config_exception("Value [" + name + "] doesn't exist or is of wrong type");
}

template <class T>
T Config_reader::read_value(const std::string &name, T default_value, const toml_lite::TablePtr &table) {
    T v;
    if (table && table->get(name, v)) {
        out << name << " = " << show(v) << std::endl;
        return v;
    }
    out << name << " = " << show(default_value) << " (default)" << std::endl;
    return default_value;
}

// src/Config_reader.cpp:18-74
void Config_reader::read_all() {
    params.species.register_species("electron", -1, 1);
    params.species.register_species("positron", 1, 1);
    params.species.register_species("proton", 1, 1836.15267343);
    params.species.register_species("photon", 0, 0, true);

    const std::string BD = "base_density_CGS", BF = "base_frequency_SI";
    if (config->contains(BD)) {
        if (config->contains(BF)) throw Config_exception("Both [" + BD + "] and [" + BF + "] exist, choose one");
        params.base_frequency_SI = density_to_frequency(read_value<double>(BD, config));
        out << "Set " << BF << " = " << params.base_frequency_SI << "\n";
    } else {
        params.base_frequency_SI = read_value<double>(BF, 1.0, config);
    }
    params.l.x = read_value<double>("lx", config);
    params.l.y = read_value<double>("ly", config);
    params.l.z = read_value<double>("lz", config);
    params.d.x = read_value<double>("dx", config);
    params.d.y = read_value<double>("dy", config);
    params.d.z = read_value<double>("dz", config);
    params.dt = read_value<double>("dt", params.l.x, config);
    params.t_end = read_value<double>("t_end", 0.0, config);
    params.ppcy = read_value<int>("ppcy", 1, config);
    params.ppcz = read_value<int>("ppcz", 1, config);
    params.magnetic_field_iterations = read_value<int>("magnetic_field_iterations", 0, config);
    params.qed = read_value<bool>("qed", false, config);
    out << std::endl;
    init_laser();
    out << std::endl;
    init_bunch();
    out << std::endl;
    init_plasma_profile();
    out << std::endl;
    init_output_parameters();
    init_device();
    init_wakes();
}

// src/Config_reader.cpp:109-136
void Config_reader::init_laser() {
    if (!config->contains("laser")) {
        out << "No [laser] block in config" << std::endl;
        return;
    }
    out << "Parsing [laser] ..." << std::endl;
    auto t = config->get_table("laser");
    if (!t) throw Config_exception("[laser] must be a table");
    parse_laser_table(t, params.laser_a0, params.laser_sigma, params.laser_r0);
    params.a_sqr = laser_a_sqr(params.laser_a0, params.laser_sigma, params.laser_r0);
    params.laser_gaussian = true;
    // extension keys: an evolving laser (no reference counterpart)
    params.laser_evolve = read_value<bool>("evolve", false, t);
    if (params.laser_evolve) {
        params.laser_k0 = read_value<double>("k0", t);
        if (!(params.laser_k0 > 0)) throw Config_exception("[laser] k0 must be positive");
        params.laser_chi_iterations = read_value<int>("chi_iterations", 1, t);
    }
}

// The Gaussian laser keys of a [laser] table (src/Config_reader.cpp:113-127).
void Config_reader::parse_laser_table(const toml_lite::TablePtr &t, double &a0, double sigma[3], double r0[3]) {
    const auto shape = read_value<std::string>("shape", t);
    if (shape != "gaussian") throw Config_exception("Laser shape \"" + shape + "\" is not supported, use \"gaussian\".");
    a0 = read_value<double>("a0", t);
    sigma[0] = read_value<double>("xsigma", t);
    sigma[1] = read_value<double>("ysigma", t);
    sigma[2] = read_value<double>("zsigma", t);
    r0[0] = read_value<double>("x0", sigma[0], t);
    r0[1] = read_value<double>("y0", 0.5 * params.l.y, t);
    r0[2] = read_value<double>("z0", 0.5 * params.l.z, t);
}

// <a^2> = a0^2/2 Gaussian with the full widths at 1/e^2 turned into Gaussian
// parameters (width / sqrt(8), src/Config_reader.cpp:128-131).
profile3d laser_a_sqr(double a0, const double sigma[3], const double r0[3]) {
    vector3d width{sigma[0], sigma[1], sigma[2]};
    width /= std::sqrt(8.0);
    return gaussian3d(0.5 * a0 * a0, width, {r0[0], r0[1], r0[2]});
}

// src/Config_reader.cpp:138-186
Bunch_parameters Config_reader::parse_bunch_table(const toml_lite::TablePtr &t) {
    Bunch_parameters b;
    b.ppc.x = read_value<int>("ppcx", 1, t);
    b.ppc.y = read_value<int>("ppcy", 1, t);
    b.ppc.z = read_value<int>("ppcz", 1, t);
    b.gamma = read_value<double>("gamma", t);
    const auto shape = read_value<std::string>("shape", t);
    if (shape == "gaussian") {
        const double rho0 = read_value<double>("rho0", t);
        const double xs = read_value<double>("xsigma", t), ys = read_value<double>("ysigma", t),
                     zs = read_value<double>("zsigma", t);
        const double x0 = read_value<double>("x0", xs, t);
        const double y0 = read_value<double>("y0", 0.5 * params.l.y, t);
        const double z0 = read_value<double>("z0", 0.5 * params.l.z, t);
        vector3d width{xs, ys, zs};
        width /= std::sqrt(8.0);
        b.rho = gaussian3d(rho0, width, {x0, y0, z0});
    } else if (shape == "parabolic" || shape == "cylindrical") {
        const double rho0 = read_value<double>("rho0", t);
        const double xsize = read_value<double>("xsize", t), rsize = read_value<double>("rsize", t);
        const double x0 = read_value<double>("x0", t);
        const double y0 = read_value<double>("y0", 0.5 * params.l.y, t);
        const double z0 = read_value<double>("z0", 0.5 * params.l.z, t);
        b.rho = shape == "parabolic" ? parabolic3d(rho0, xsize, rsize, {x0, y0, z0})
                                     : cylindrical3d(rho0, xsize, rsize, {x0, y0, z0});
    } else {
        throw Config_exception("Bunch shape \"" + shape +
                               "\" is not supported, use \"gaussian\", \"parabolic\", or \"cylindrical\".");
    }
    const std::string sp = read_value<std::string>("species", std::string("electron"), t);
    b.species_id = params.species.find_species(sp);
    return b;
}

// src/Config_reader.cpp:188-216
void Config_reader::init_bunch() {
    if (config->is_table("bunch")) {
        out << "Parsing [bunch] ..." << std::endl;
        params.bunch_parameters_array.push_back(parse_bunch_table(config->get_table("bunch")));
    } else if (config->is_table_array("bunch")) {
        out << "Parsing [[bunch]] ..." << std::endl;
        const auto *arr = config->get_table_array("bunch");
        out << "Size " << arr->size() << std::endl;
        for (size_t i = 0; i < arr->size(); ++i) {
            out << "Parsing element " << i << "..." << std::endl;
            params.bunch_parameters_array.push_back(parse_bunch_table((*arr)[i]));
        }
    } else if (config->contains("bunch")) {
        out << "\"bunch\" cannot be an element, only a table or an array of tables" << std::endl;
    } else {
        out << "No [bunch] or [[bunch]] in config" << std::endl;
    }
}

// src/Config_reader.cpp:218-234
void Config_reader::init_output_parameters() {
    Output_parameters o;
    if (config->contains("output")) {
        out << "Parsing [output] ..." << std::endl;
        auto t = config->get_table("output");
        o.output3d = read_value<bool>("fields_3d", false, t);
        o.output_xy = read_value<bool>("fields_xy", false, t);
        o.output_bunch = read_value<bool>("bunch", false, t);
        o.z0 = read_value<double>("z0", 0.5 * params.l.z, t);
    } else {
        out << "No [output] block in config; no output files will be created" << std::endl;
    }
    params.output_parameters = o;
}

// src/Config_reader.cpp:236-258, plus the "channel" and "ramp" extension profiles
void Config_reader::init_plasma_profile() {
    if (!config->contains("plasma")) {
        out << "No [plasma] block in config; uniform plasma will be used";
        return;
    }
    out << "Parsing [plasma] ..." << std::endl;
    params.plasma_profile = parse_plasma_table(config->get_table("plasma"));
}

profile2d Config_reader::parse_plasma_table(const toml_lite::TablePtr &t) {
    if (!t) throw Config_exception("[plasma] must be a table");
    const auto profile = read_value<std::string>("profile", t);
    if (profile == "powerlaw") {
        const double power = read_value<double>("power", t);
        const double radius = read_value<double>("radius", t);
        const double mn = read_value<double>("min_factor", 0.0, t);
        const double mx = read_value<double>("max_factor", 1e10, t);
        const double y0 = read_value<double>("y0", 0.5 * params.l.y, t);
        const double z0 = read_value<double>("z0", 0.5 * params.l.z, t);
        return powerlaw2d(power, radius, y0, z0, mn, mx, -1.0);
    }
    if (profile == "constant") return constant2d(read_value<double>("rho0", -1.0, t));
    if (profile == "channel") {
        const double base = read_value<double>("base", 1.0, t);
        const double curv = read_value<double>("curvature", t);
        const double y0 = read_value<double>("y0", 0.5 * params.l.y, t);
        const double z0 = read_value<double>("z0", 0.5 * params.l.z, t);
        return channel2d(base, curv, y0, z0, -1.0);
    }
    if (profile == "ramp") {
        const double slope = read_value<double>("slope", t);
        const double y0 = read_value<double>("y0", 0.5 * params.l.y, t);
        return ramp2d(slope, y0, -1.0);
    }
    throw Config_exception("Plasma profile \"" + profile +
                           "\" is not supported, use \"powerlaw\", \"constant\", \"channel\" or \"ramp\".");
}

namespace {
// The member's own sub-table over the deck's top-level table of that name:
// keys the member leaves out are taken from the deck.
toml_lite::TablePtr overlay(const toml_lite::TablePtr &base, const toml_lite::TablePtr &member) {
    auto t = std::make_shared<toml_lite::Table>();
    if (base) t->values = base->values;
    if (member)
        for (const auto &kv : member->values) t->values[kv.first] = kv.second;
    return t;
}
}  // namespace

// [[wake]]: a batch of independent wakes on the deck's geometry (extension, no
// reference counterpart).  Each member may carry name, ppcy, ppcz and its own
// [wake.laser] / [wake.plasma] tables, whose keys override the deck's [laser] /
// [plasma].
void Config_reader::init_wakes() {
    if (!config->contains("wake")) return;
    if (!config->is_table_array("wake")) throw Config_exception("\"wake\" must be an array of tables [[wake]]");
    if (!params.bunch_parameters_array.empty()) throw Config_exception("[[wake]] batches take no [bunch]");
    if (params.laser_evolve) throw Config_exception("[[wake]] batches take no evolving laser");
    if (params.device.precision != 1)
        throw Config_exception("[[wake]] batches run on the mixed path: [device] precision must be \"mixed\"");
    const auto *arr = config->get_table_array("wake");
    out << "Parsing [[wake]] ..." << std::endl << "Size " << arr->size() << std::endl;
    const auto laser = config->is_table("laser") ? config->get_table("laser") : nullptr;
    const auto plasma = config->is_table("plasma") ? config->get_table("plasma") : nullptr;
    for (size_t i = 0; i < arr->size(); ++i) {
        const auto &m = (*arr)[i];
        out << "Parsing element " << i << "..." << std::endl;
        Wake_member w;
        w.name = read_value<std::string>("name", "wake" + std::to_string(i), m);
        if (w.name.empty() || w.name.find_first_not_of("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
                                                        "0123456789_-.") != std::string::npos)
            throw Config_exception("[[wake]] name \"" + w.name + "\" must be letters, digits, '_', '-' or '.'");
        for (const auto &o : params.wakes)
            if (o.name == w.name) throw Config_exception("[[wake]] name \"" + w.name + "\" is used twice");
        w.ppcy = read_value<int>("ppcy", params.ppcy, m);
        w.ppcz = read_value<int>("ppcz", params.ppcz, m);
        if (w.ppcy <= 0 || w.ppcz <= 0) throw Config_exception("[[wake]] ppcy and ppcz must be positive");
        const auto ml = m->is_table("laser") ? m->get_table("laser") : nullptr;
        if (!laser && !ml) throw Config_exception("[[wake]] member " + w.name + " has no laser");
        const auto lt = overlay(laser, ml);
        parse_laser_table(lt, w.a0, w.sigma, w.r0);
        bool evolve = false;
        if (lt->get("evolve", evolve) && evolve) throw Config_exception("[[wake]] batches take no evolving laser");
        const auto mp = m->is_table("plasma") ? m->get_table("plasma") : nullptr;
        w.plasma_profile = (mp || plasma) ? parse_plasma_table(overlay(plasma, mp)) : constant2d(-1.0);
        params.wakes.push_back(w);
    }
}

// [device] table: options of this build (no reference counterpart).
void Config_reader::init_device() {
    if (!config->contains("device")) return;
    out << "Parsing [device] ..." << std::endl;
    auto t = config->get_table("device");
    const auto prec = read_value<std::string>("precision", std::string("mixed"), t);
    if (prec == "mixed") params.device.precision = 1;
    else if (prec == "fp64") params.device.precision = 0;
    else throw Config_exception("[device] precision must be \"mixed\" or \"fp64\"");
    params.device.device = read_value<int>("device", 0, t);
    params.device.use_graph = read_value<bool>("graph", true, t) ? 1 : 0;
    params.device.resort_every = read_value<int>("resort_every", 0, t);
    params.device.thermal_momentum = read_value<double>("thermal_momentum", 0.0, t);
    params.device.seed = (unsigned long long)read_value<int>("seed", 1234, t);
    params.device.stream_output = read_value<bool>("stream_output", true, t) ? 1 : 0;
    const auto bunch = read_value<std::string>("bunch", std::string("device"), t);
    if (bunch == "device") params.device.bunch_on_device = 1;
    else if (bunch == "host") params.device.bunch_on_device = 0;
    else throw Config_exception("[device] bunch must be \"device\" or \"host\"");
}
