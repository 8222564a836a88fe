// host_selftest -- CPU checks of the host layer (no GPU needed):
// TOML subset, Config_reader on the reference's input decks (keys, defaults,
// errors), Species_register, profiles, HDF5 layout round trip.
// Prints "PASS <name>" / "FAIL <name>: why" per check; exit code = failures.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "Config_reader.h"
#include "Output_writer.h"
#include "Species_register.h"
#include "toml_lite.h"

static int failures = 0;
static void check(bool ok, const std::string &name, const std::string &why = "") {
    if (ok) {
        std::cout << "PASS " << name << "\n";
    } else {
        std::cout << "FAIL " << name << ": " << why << "\n";
        ++failures;
    }
}
static bool near(double a, double b, double tol = 1e-12) { return std::fabs(a - b) <= tol * (1 + std::fabs(b)); }

template <class F>
static bool throws(F f) {
    try {
        f();
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

int main(int argc, char **argv) {
    const std::string decks = argc > 1 ? argv[1] : "tests/golden/decks";
    const std::string tmp = argc > 2 ? argv[2] : "/tmp";
    std::ostringstream sink;

    // TOML subset
    {
        auto t = toml_lite::parse_string("a = 1\nb = 2.5 # c\nc = \"x#y\"\nd = true\n[t]\ne = -3e2\n[[arr]]\nf = 1\n[[arr]]\nf = 2\n");
        double a = 0, b = 0, e = 0;
        int ai = 0;
        std::string c;
        bool dv = false;
        check(t->get("a", a) && a == 1.0 && t->get("a", ai) && ai == 1, "toml int readable as double and int");
        check(t->get("b", b) && b == 2.5 && !t->get("b", ai), "toml float not readable as int");
        check(t->get("c", c) && c == "x#y", "toml string with hash");
        check(t->get("d", dv) && dv, "toml bool");
        check(t->get_table("t") && t->get_table("t")->get("e", e) && e == -300.0, "toml table");
        check(t->get_table_array("arr") && t->get_table_array("arr")->size() == 2, "toml array of tables");
        check(throws([] { toml_lite::parse_string("a = 1\na = 2\n"); }), "toml duplicate key rejected");
        check(throws([] { toml_lite::parse_string("a = \"x\n"); }), "toml unterminated string rejected");
    }

    // Config_reader on the reference's decks (inputs/test/*.toml)
    {
        System_parameters p = Config_reader(decks + "/linear_wake_test.toml", sink).get_parameters();
        check(p.l.x == 60 && p.l.y == 20 && p.d.x == 0.1 && p.magnetic_field_iterations == 5, "linear_wake keys");
        check(p.dt == 60 && p.t_end == 0 && p.ppcy == 1 && !p.qed && p.base_frequency_SI == 1.0,
              "linear_wake defaults (dt = lx, t_end = 0)");
        const double w = 4.0 / std::sqrt(8.0), wr = 4.0 / std::sqrt(8.0);
        const double want = 0.5 * 0.01 * 0.01 * std::exp(-(44.0 - 45.0) * (44.0 - 45.0) / w / w - 0.25 / wr / wr);
        check(near(p.a_sqr(44.0, 10.5, 10.0), want), "laser a_sqr = a0^2/2 gauss with widths / sqrt(8)");
        check(p.laser_gaussian && p.laser_r0[0] == 45 && p.laser_r0[1] == 10, "laser x0, default y0 = ly/2");
        check(p.output_parameters.output3d && p.output_parameters.output_xy && !p.output_parameters.output_bunch &&
                  p.output_parameters.output_photons && p.output_parameters.z0 == 10,
              "output flags and z0 default");
        check(p.plasma_profile(1.0, 2.0) == -1.0, "default plasma constant -1");
    }
    {
        System_parameters p = Config_reader(decks + "/betatron_oscillations.toml", sink).get_parameters();
        check(p.bunch_parameters_array.size() == 1, "bunch table parsed");
        const auto &b = p.bunch_parameters_array[0];
        check(b.gamma == 800 && b.species_id == 0 && b.ppc.x == 1, "bunch gamma, species electron, ppc default");
        check(near(b.rho(10.0, 5.0, 5.0), 2.0) && b.rho(10.0, 6.5, 5.0) == 0.0, "parabolic bunch profile");
        check(p.dt == 6.3 && p.t_end == 252, "dt and t_end read");
        check(std::fabs(p.base_frequency_SI - 5.641460e13) / 5.641460e13 < 1e-5, "base_density_CGS -> frequency");
    }
    {
        System_parameters p = Config_reader(decks + "/profile_test.toml", sink).get_parameters();
        check(near(p.plasma_profile(10.0 + 3.0, 10.0 + 4.0), -(5.0 / 20.0) * (5.0 / 20.0)), "powerlaw profile");
    }
    {
        System_parameters p = Config_reader(decks + "/bunch_test.toml", sink).get_parameters();
        const auto &b = p.bunch_parameters_array[0];
        check(near(b.rho(15.0, 10.0, 10.0), 3.0), "gaussian bunch peak");
    }
    check(throws([&] { Config_reader::from_string("base_density_CGS = 1e18\nbase_frequency_SI = 1\nlx=1\n", sink); }),
          "both base_density_CGS and base_frequency_SI rejected");
    check(throws([&] { Config_reader::from_string("ly = 1\nlz = 1\ndx = 1\ndy = 1\ndz = 1\n", sink); }),
          "missing lx rejected");
    check(throws([&] { Config_reader::from_string("lx = \"a\"\nly = 1\nlz = 1\ndx = 1\ndy = 1\ndz = 1\n", sink); }),
          "wrongly typed lx rejected");
    check(throws([&] {
              Config_reader::from_string("lx=1\nly=1\nlz=1\ndx=1\ndy=1\ndz=1\n[laser]\nshape='flat'\n", sink);
          }),
          "unknown laser shape rejected");
    // a wrongly typed key that has a default falls back to the default, as the
    // reference's read_value(name, default) does (src/Config_reader.cpp:97-107)
    check(Config_reader::from_string("lx=1\nly=1\nlz=1\ndx=1\ndy=1\ndz=1\nppcy = 1.5\n", sink).get_parameters().ppcy == 1,
          "float ppcy falls back to the default");
    {
        System_parameters p = Config_reader::from_string(
                                  "lx=4\nly=10\nlz=10\ndx=0.1\ndy=0.5\ndz=0.5\n[plasma]\nprofile='channel'\ncurvature=0.05\n"
                                  "[device]\nprecision='fp64'\nresort_every=4\nthermal_momentum=0.001\n",
                                  sink)
                                  .get_parameters();
        check(near(p.plasma_profile(5.0 + 3.0, 5.0 + 4.0), -(1.0 + 0.05 * 25.0)), "channel profile extension");
        check(p.device.precision == 0 && p.device.resort_every == 4 && p.device.thermal_momentum == 0.001,
              "device table extension");
    }

    // dotted headers: [a.b] inside a table and inside the last [[a]] element
    {
        auto t = toml_lite::parse_string("[u]\nx = 1\n[u.v]\ny = 2\n[[w]]\nn = 'a'\n[w.l]\nk = 3\n[[w]]\nn = 'b'\n"
                                         "[w.l]\nk = 4\n");
        int y = 0, k0 = 0, k1 = 0;
        check(t->get_table("u")->get_table("v")->get("y", y) && y == 2, "toml dotted table in a table");
        const auto *w = t->get_table_array("w");
        check(w && w->size() == 2 && (*w)[0]->get_table("l")->get("k", k0) && k0 == 3 &&
                  (*w)[1]->get_table("l")->get("k", k1) && k1 == 4,
              "toml dotted table in the last array element");
        check(throws([] { toml_lite::parse_string("[u]\n[u.v]\n[u.v]\n"); }), "toml duplicate dotted table rejected");
        check(throws([] { toml_lite::parse_string("x = 1\n[x.v]\n"); }), "toml dotted table over a value rejected");
    }

    // [[wake]] batch members (extension): inheritance from [laser] / [plasma],
    // per-member overrides, names and the combinations a batch cannot run
    {
        const std::string base = "lx=3\nly=20\nlz=20\ndx=0.1\ndy=0.625\ndz=0.625\nppcy=2\nppcz=2\n"
                                 "[laser]\nshape='gaussian'\na0=1.0\nxsigma=4\nysigma=6\nzsigma=6\nx0=1.5\n";
        System_parameters p =
            Config_reader::from_string(base + "[[wake]]\nname='low'\n[[wake]]\nppcy=1\n[wake.laser]\na0=2.5\ny0=9\n"
                                              "[wake.plasma]\nprofile='channel'\ncurvature=0.05\n",
                                       sink)
                .get_parameters();
        check(p.wakes.size() == 2 && p.wakes[0].name == "low" && p.wakes[1].name == "wake1", "wake names and default");
        const auto &a = p.wakes[0], &b = p.wakes[1];
        check(a.a0 == 1.0 && a.sigma[1] == 6 && a.r0[0] == 1.5 && a.r0[1] == 10 && a.ppcy == 2,
              "wake member inherits [laser] and ppc");
        check(b.a0 == 2.5 && b.r0[1] == 9 && b.r0[2] == 10 && b.sigma[0] == 4 && b.ppcy == 1 && b.ppcz == 2,
              "wake member overrides a0, y0 and ppcy");
        check(a.plasma_profile(3.0, 4.0) == -1.0 && near(b.plasma_profile(10.0 + 3.0, 10.0 + 4.0), -(1.0 + 0.05 * 25.0)),
              "wake member plasma profile");
        check(throws([&] { Config_reader::from_string(base + "[[wake]]\nname='a'\n[[wake]]\nname='a'\n", sink); }),
              "duplicate wake name rejected");
        check(throws([&] { Config_reader::from_string(base + "[[wake]]\nname='a/b'\n", sink); }),
              "wake name with a path separator rejected");
        check(throws([&] { Config_reader::from_string(base + "[wake]\nname='a'\n", sink); }),
              "[wake] as a single table rejected");
        check(throws([&] { Config_reader::from_string(base + "[device]\nprecision='fp64'\n[[wake]]\n", sink); }),
              "wake batch on the fp64 path rejected");
        check(throws([&] {
                  Config_reader::from_string(base + "[bunch]\nshape='gaussian'\ngamma=10\nrho0=1\nxsigma=1\n"
                                                    "ysigma=1\nzsigma=1\n[[wake]]\n",
                                             sink);
              }),
              "wake batch with a bunch rejected");
        check(throws([&] {
                  Config_reader::from_string("lx=3\nly=20\nlz=20\ndx=0.1\ndy=0.625\ndz=0.625\n[[wake]]\n", sink);
              }),
              "wake member without a laser rejected");
    }

    // Species_register (src/Species_register.cpp:6-30)
    {
        Species_register s;
        check(s.register_species("e", -1, 1) == 0 && s.register_species("p", 1, 1836) == 1, "species ids");
        check(throws([&] { s.register_species("e", 1, 1); }), "duplicate species rejected");
        check(s.find_species("x") == -1 && s.is_registered("p") && s[1].charge_to_mass_ratio == 1.0 / 1836,
              "species lookup");
        s.register_species("photon", 0, 0, true);
        check(s[2].charge_to_mass_ratio == 0.0 && s[2].is_photon, "massless species");
    }

    // HDF5 layout round trip
    {
        Output_parameters op;
        op.output3d = true;
        op.output_xy = true;
        op.output_bunch = true;
        op.output_photons = false;
        op.z0 = 0.75;
        // Output_writer writes into the working directory, like the reference
        if (chdir(tmp.c_str()) != 0) check(false, "chdir", tmp);
        ivector3d n{3, 4, 2};
        vector3d d{0.5, 0.25, 0.5};
        {
            Output_writer w(op, 7);
            array2d psi({4, 2}, d, {0, 0, 0}, Plane::YZ);
            w.initialize_slice_array(n, d, psi, "psi");
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j < 4; ++j)
                    for (int k = 0; k < 2; ++k) psi(j, k) = 100 * i + 10 * j + k;
                w.write_slice(psi, "psi", i);
            }
            array3d ex(n, d, {-0.25, 0, 0});
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 4; ++j)
                    for (int k = 0; k < 2; ++k) ex(i, j, k) = i + j + 0.5 * k;
            w.write_array(ex, "ex");
            std::vector<bunch_particle_3d> b(2);
            b[0].x = 1.5;
            b[1].n = 2.0;
            w.write_bunch_parameters(b, 3.0);
        }
        int rank = 0;
        unsigned long long dims[4];
        check(qpic_h5_dataset_shape("Fields_007.h5", "psi", &rank, dims) == 0 && rank == 3 && dims[0] == 3 &&
                  dims[1] == 4 && dims[2] == 2,
              "Fields_NNN.h5 slice dataset [nx, ny, nz]");
        std::vector<double> v(24);
        check(qpic_h5_read_dataset("Fields_007.h5", "psi", v.data(), 24) == 0 && v[((2 * 4) + 3) * 2 + 1] == 231.0,
              "slice hyperslab [i, j, k]");
        double o[3] = {9, 9, 9};
        check(qpic_h5_read_attr("Fields_007.h5", "psi", "origin", o, 3) == 0 && o[0] == 0 && o[1] == 0 && o[2] == 0,
              "slice array origin attribute");
        double st[3] = {0, 0, 0};
        check(qpic_h5_read_attr("Fields_007.h5", "psi", "steps", st, 3) == 0 && st[0] == 0.5 && st[1] == 0.25,
              "slice array steps attribute");
        check(qpic_h5_read_attr("Fields_007.h5", "ex", "origin", o, 3) == 0 && o[0] == -0.25,
              "3D array origin attribute (e_x staggering)");
        std::vector<double> xy(12);
        // z0 = 0.75 -> k0 = 1, frac 0.5 between k = 1 and the clamped last node
        check(qpic_h5_read_dataset("Fields_xy_007.h5", "psi", xy.data(), 12) == 0 && xy[1 * 4 + 2] == 121.0,
              "xy line interpolation at z0");
        int plane = 0;
        check(qpic_h5_read_attr_int("Fields_xy_007.h5", "psi", "plane", &plane) == 0 && plane == 1,
              "plane attribute XY");
        double pc = 0;
        check(qpic_h5_read_attr("Fields_xy_007.h5", "psi", "plane_coordinate", &pc, 1) == 0 && pc == 0.75,
              "plane_coordinate attribute");
        check(qpic_h5_read_dataset("Fields_xy_007.h5", "ex", xy.data(), 12) == 0 && xy[2 * 4 + 1] == 2 + 1 + 0.5,
              "xy plane of a 3D array");
        double w[2];
        check(qpic_h5_read_dataset("Bunch_007.h5", "weight", w, 2) == 0 && w[1] == 6.0, "bunch weight = n * norm");
        check(qpic_h5_read_dataset("Bunch_007.h5", "x", w, 2) == 0 && w[0] == 1.5, "bunch x dataset");
    }
    std::cout << (failures ? "FAILED " : "ALL PASSED ") << failures << std::endl;
    return failures;
}
