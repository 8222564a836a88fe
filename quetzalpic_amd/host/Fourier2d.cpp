#include "Fourier2d.h"

#include <cstdlib>
#include <stdexcept>
#include <string>

static void check(int st, const char *what) {
    if (st < 0) throw std::runtime_error(std::string(what) + ": " + qp_last_error());
}

qp_ctx qp_host_context() {
    static qp_ctx ctx = nullptr;
    if (!ctx) {
        const char *dev = std::getenv("QP_DEVICE");
        check(qp_ctx_create(dev ? std::atoi(dev) : 0, &ctx), "qp_ctx_create");
    }
    return ctx;
}

Fourier2d::Fourier2d(int n1_, int n2_) : n1(n1_), n2(n2_) {
    check(qp_fourier2d_create(qp_host_context(), n1, n2, &handle), "qp_fourier2d_create");
    in = new double[(size_t)n1 * n2]();
    out = new double[(size_t)n1 * (n2 / 2 + 1)][2]();
}

void Fourier2d::release() {
    if (handle) qp_fourier2d_destroy(handle);
    delete[] in;
    delete[] out;
    handle = nullptr;
    in = nullptr;
    out = nullptr;
}

Fourier2d::~Fourier2d() { release(); }

Fourier2d &Fourier2d::operator=(Fourier2d &&o) noexcept {
    if (this != &o) {
        release();
        n1 = o.n1;
        n2 = o.n2;
        handle = o.handle;
        in = o.in;
        out = o.out;
        o.handle = nullptr;
        o.in = nullptr;
        o.out = nullptr;
    }
    return *this;
}

void Fourier2d::forward_transform() { check(qp_fourier2d_forward(handle, in, &out[0][0]), "forward_transform"); }

void Fourier2d::backward_transform() { check(qp_fourier2d_backward(handle, &out[0][0], in), "backward_transform"); }

void Fourier2d::solve_poisson(double d1, double d2, double D) {
    check(qp_solve2d(handle, d1, d2, D, in, in), "qp_solve2d");
}
