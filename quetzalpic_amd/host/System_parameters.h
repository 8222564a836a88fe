// System_parameters.h -- run parameters (the fields of the reference's
// src/System_parameters.h:8-41, plus the device-path knobs of this build).
#pragma once

#include <string>
#include <vector>

#include "Species_register.h"
#include "containers.h"
#include "profiles.h"

struct Output_parameters {
    bool output3d = false;
    bool output_xy = false;
    bool output_bunch = false;
    bool output_photons = true;
    double z0 = 0;
};

struct Bunch_parameters {
    profile3d rho = constant3d(0.0);
    ivector3d ppc = {1, 1, 1};
    double gamma = 1.0;
    int species_id = 0;
};

// Device-path options (no reference counterpart).
struct Device_parameters {
    int device = 0;
    int precision = 1;      // 0 = plain FP64 path, 1 = fused mixed path
    int use_graph = 1;
    int resort_every = 0;
    double thermal_momentum = 0.0;  // extension: N(0, sigma) initial p_y, p_z
    unsigned long long seed = 1234;
    int bunch_on_device = 1;        // driver bunch in qp_bunch (without QED) or on the host
    int stream_output = 1;          // mixed path: per-slice output streamed to a writer thread
};

// One member of a [[wake]] batch (extension): its name (output file prefix),
// lattice, plasma profile and Gaussian laser.
struct Wake_member {
    std::string name;
    int ppcy = 1, ppcz = 1;
    profile2d plasma_profile = constant2d(-1.0);
    double a0 = 0, sigma[3] = {0, 0, 0}, r0[3] = {0, 0, 0};
};

struct System_parameters {
    vector3d l{0, 0, 0};
    vector3d d{0, 0, 0};
    double dt = 0;
    double t_end = 0;
    int ppcy = 1;
    int ppcz = 1;
    int magnetic_field_iterations = 1;
    profile3d a_sqr = constant3d(0.0);
    profile2d plasma_profile = constant2d(-1.0);
    // Gaussian laser parameters when the laser is Gaussian (device evaluation)
    bool laser_gaussian = false;
    double laser_a0 = 0, laser_sigma[3] = {0, 0, 0}, laser_r0[3] = {0, 0, 0};
    // Extension: the laser envelope evolves between beam steps ([laser] evolve, k0, chi_iterations)
    bool laser_evolve = false;
    double laser_k0 = 0;
    int laser_chi_iterations = 1;
    Output_parameters output_parameters;
    std::vector<Bunch_parameters> bunch_parameters_array;
    Species_register species;
    double base_frequency_SI = 1;
    bool qed = false;
    Device_parameters device;
    std::vector<Wake_member> wakes;  // [[wake]] batch; empty for a single System_3d
};
