// System_3d.h -- the reference's simulation driver interface
// (src/System_3d.h:24-99) with the slice loop on the MI355X: solve_wakefield
// hands the plasma to qp_wake_run (include/quetzal_hip.h) and fills the host
// 3D arrays and HDF5 output from the per-slice callback.  The driver-bunch
// push (src/System_3d.cpp:199-236, 750-946) stays on the host.
#pragma once

#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "../../include/quetzal_hip.h"
#include "Output_writer.h"
#include "Plasma_units.h"
#include "System_parameters.h"
#include "arrays.h"

class System_3d {
public:
    System_3d(System_parameters &params, std::ostream &out);
    ~System_3d();
    System_3d(const System_3d &) = delete;
    System_3d &operator=(const System_3d &) = delete;
    void run();

    // Access for drivers and tests: the 3D arrays by reference name
    // ("aSqr", "susceptibility", "rho_bunch", "jx_bunch", "ex", "ey", "ez", "by", "bz").
    const array3d &field3d(const std::string &name) const;
    ivector3d dimensions() const { return n; }
    vector3d steps() const { return d; }
    int stopped_slice() const { return stop_slice_; }
    double last_wake_ms() const { return last_wake_ms_; }

private:
    void solve_wakefield(int iteration, Output_writer &output_writer);
    size_t count_wake_particles() const;
    void init_wake_particles();
    size_t count_bunch_particles(ivector3d ppc, const profile3d &rho) const;
    void init_bunch_particles(size_t index, const Bunch_parameters &bunch);
    void init_a_sqr(const profile3d &func);
    void deposit_wakefield_to_bunch();
    void advance_bunch_particles(std::vector<bunch_particle_3d> &photons, double &max_radiation_probability);
    void deposit_bunch_densities();
    static int slice_callback(const qp_slice_view *v, void *user);

    ivector3d n{0, 0, 0};
    vector3d l{0, 0, 0};
    vector3d d{0, 0, 0};
    double dt = 0, t_end = 0;
    int time_iterations = 0;
    int ppcy = 1, ppcz = 1;
    profile2d plasma_profile;
    int magnetic_field_iterations = 0;
    double psi_threshold = 1e-4;
    double magnetic_field_D = 10.0;
    Plasma_units plasma_units;
    Output_parameters output_parameters;
    Device_parameters device;
    std::ostream &out;

    std::vector<double> wy, wz, wpy, wpz, wn;  // wake particles, host SoA
    std::vector<bunch_particle_3d> bunch_particles;

    array3d a_sqr, susceptibility, rho_bunch, jx_bunch, ex, ey, ez, by, bz;
    array2d psi, rho, jx, jy, jz;

    Species_register species;
    bool qed = false;
    std::mt19937_64 qed_rng;

    qp_wake wake = nullptr;
    qp_bunch dbunch = nullptr;  // driver bunch on the device (no QED), attached to the wake
    qp_envelope env = nullptr;  // evolving laser envelope ([laser] evolve), attached to the wake
    int env_chi_iterations = 1;
    void upload_bunch();
    void download_bunch();
    Output_writer *cur_writer = nullptr;
    int stop_slice_ = -1;
    double last_wake_ms_ = 0;
};
