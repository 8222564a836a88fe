// Fourier2d.h -- the reference's FFTW wrapper interface (src/Fourier2d.h:4-19)
// on the MI355X: public host buffers `in` (n1*n2 real) and `out`
// (n1*(n2/2+1) complex, re/im pairs), unnormalised r2c / c2r executed on the
// device through the C-ABI (qp_fourier2d_*).
#pragma once

#include "../../include/quetzal_hip.h"

class Fourier2d {
public:
    double *in = nullptr;
    double (*out)[2] = nullptr;
    Fourier2d() = default;
    Fourier2d(int n1, int n2);
    ~Fourier2d();
    Fourier2d(const Fourier2d &) = delete;
    Fourier2d &operator=(const Fourier2d &) = delete;
    Fourier2d &operator=(Fourier2d &&other) noexcept;
    void forward_transform();   // in -> out
    void backward_transform();  // out -> in
    // solve_poisson_equation(D) of src/System_3d.cpp:686-706 on `in`, in place
    // (unnormalised, like the reference; the caller divides by n1*n2).
    void solve_poisson(double d1, double d2, double D);

private:
    int n1 = 0, n2 = 0;
    qp_fourier2d handle = nullptr;
    void release();
};

// Process-wide device context of the host layer (device from $QP_DEVICE, default 0).
qp_ctx qp_host_context();
