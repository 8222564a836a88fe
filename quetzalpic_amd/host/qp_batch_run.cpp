// qp_batch_run -- run a batch of independent wakes through the C-ABI from a
// member list, for timing and for profilers that need a plain native process.
//
//   qp_batch_run <members-file> [passes] [graph 0|1] [slices]
//
// `slices` runs only the first slices of the geometry (the first of the pass,
// from the far end of the box): the same per-slice work, for short profiler
// runs.
// Members file (text, '#' comments), as written by
// `python -m quetzalpic_amd.workloads members ...`:
//   geometry nx ny nz dx dy dz ly lz K
//   member a0 xsigma ysigma zsigma x0 y0 z0 ppc kind p1 p2
// with kind constant (n = -p1), channel (n = -(p1 + p2 r^2)) or ramp
// (n = -max(0, 1 + p1 (y - ly/2))).  Particles follow init_wake_particles
// (src/System_3d.cpp:724-748): lattice ((i+1/2) dy/ppc, (j+1/2) dz/ppc), i
// outer, weight = profile / ppc^2, cells with zero profile skipped.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/quetzal_hip.h"

namespace {

struct Member {
    double a0, xs, ys, zs, x0, y0, z0;
    int ppc;
    std::string kind;
    double p1, p2;
};

double profile(const Member &m, double y, double z, double ly, double lz) {
    if (m.kind == "constant") return -m.p1;
    if (m.kind == "channel") {
        const double dy = y - ly / 2, dz = z - lz / 2;
        return -(m.p1 + m.p2 * (dy * dy + dz * dz));
    }
    if (m.kind == "ramp") return -std::max(0.0, 1.0 + m.p1 * (y - ly / 2));
    throw std::runtime_error("unknown profile kind " + m.kind);
}

void check(int st, const char *what) {
    if (st < 0) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, st, qp_last_error());
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <members-file> [passes] [graph 0|1] [slices]\n", argv[0]);
        return 2;
    }
    const int passes = argc > 2 ? std::atoi(argv[2]) : 1;
    const int graph = argc > 3 ? std::atoi(argv[3]) : 1;
    const int max_slices = argc > 4 ? std::atoi(argv[4]) : 0;
    std::ifstream in(argv[1]);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    qp_wake_params p{};
    bool have_geom = false;
    std::vector<Member> mem;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "geometry") {
            ss >> p.nx >> p.ny >> p.nz >> p.dx >> p.dy >> p.dz >> p.ly >> p.lz >> p.K;
            have_geom = true;
        } else if (tag == "member") {
            Member m;
            ss >> m.a0 >> m.xs >> m.ys >> m.zs >> m.x0 >> m.y0 >> m.z0 >> m.ppc >> m.kind >> m.p1 >> m.p2;
            if (!ss) {
                std::fprintf(stderr, "bad member line: %s\n", line.c_str());
                return 2;
            }
            mem.push_back(m);
        }
    }
    if (!have_geom || mem.empty()) {
        std::fprintf(stderr, "members file needs a geometry line and at least one member\n");
        return 2;
    }
    if (max_slices > 0 && max_slices < p.nx) p.nx = max_slices;
    p.D = 10.0;
    p.psi_threshold = 1e-4;
    p.b_threshold = 100.0;
    p.precision = QP_PREC_MIXED;
    p.use_graph = graph;
    p.resort_every = 0;

    qp_ctx ctx = nullptr;
    check(qp_ctx_create(0, &ctx), "qp_ctx_create");
    qp_wake b = nullptr;
    check(qp_wake_create_batch(ctx, &p, (int)mem.size(), &b), "qp_wake_create_batch");
    size_t total = 0;
    for (size_t i = 0; i < mem.size(); ++i) {
        const Member &m = mem[i];
        std::vector<double> y, z, w;
        for (int a = 0; a < m.ppc * p.ny; ++a) {
            for (int c = 0; c < m.ppc * p.nz; ++c) {
                const double yy = (a + 0.5) * p.dy / m.ppc, zz = (c + 0.5) * p.dz / m.ppc;
                const double v = profile(m, yy, zz, p.ly, p.lz);
                if (v == 0.0) continue;
                y.push_back(yy);
                z.push_back(zz);
                w.push_back(v / m.ppc / m.ppc);
            }
        }
        total += y.size();
        check(qp_wake_set_particles_at(b, (int)i, y.size(), y.data(), z.data(), nullptr, nullptr, w.data()),
              "qp_wake_set_particles_at");
        check(qp_wake_set_laser_gaussian_at(b, (int)i, m.a0, m.xs, m.ys, m.zs, m.x0, m.y0, m.z0),
              "qp_wake_set_laser_gaussian_at");
    }
    std::vector<int> stops(mem.size(), -1);
    for (int r = 0; r < passes; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        check(qp_wake_run_batch(b, nullptr, nullptr, stops.data()), "qp_wake_run_batch");
        const auto t1 = std::chrono::steady_clock::now();
        double dev = 0;
        check(qp_wake_last_time_ms(b, &dev), "qp_wake_last_time_ms");
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        int nstop = 0;
        for (int s : stops) nstop += s >= 0 ? 1 : 0;
        std::printf("pass %d: %zu members, %zu particles, %d slices: %.2f ms (device %.2f ms), %.3e pushes/s, "
                    "%d stopped\n",
                    r, mem.size(), total, p.nx, ms, dev, (double)total * p.nx / (ms * 1e-3), nstop);
        std::fflush(stdout);
    }
    qp_wake_destroy(b);
    qp_ctx_destroy(ctx);
    return 0;
}
