// Wake_batch.h -- a [[wake]] deck run as one batch of independent wakes
// (qp_wake_create_batch, include/quetzal_hip.h).  No reference counterpart:
// the reference runs one System_3d per process (src/quasistatic_pic.cpp:8-47).
// Every member shares the deck's geometry, time steps and magnetic-field
// iterations, and has its own lattice, plasma profile and Gaussian laser.  Each
// writes the reference's file layout under its own prefix (<name>_Fields_NNN.h5
// ...), with what System_3d writes for the same member run alone without a bunch.
#pragma once

#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/quetzal_hip.h"
#include "Output_writer.h"
#include "System_parameters.h"
#include "arrays.h"

class Wake_batch {
public:
    Wake_batch(System_parameters &params, std::ostream &out);
    ~Wake_batch();
    Wake_batch(const Wake_batch &) = delete;
    Wake_batch &operator=(const Wake_batch &) = delete;
    void run();

    int size() const { return (int)members.size(); }
    // stop slice of the member in the last pass, or -1
    int stopped_slice(int member) const { return members.at(member).stop; }
    double last_wake_ms() const { return last_wake_ms_; }

private:
    struct Member {
        Wake_member p;
        size_t particles = 0;
        int stop = -1;
        std::unique_ptr<Output_writer> writer;
        array3d a_sqr, susceptibility, rho_bunch, jx_bunch, ex, ey, ez, by, bz;
        array2d psi, rho, jx, jy, jz;
    };
    static int slice_callback(int member, const qp_slice_view *v, void *user);
    void init_member(int index);

    ivector3d n{0, 0, 0};
    vector3d l{0, 0, 0};
    vector3d d{0, 0, 0};
    double dt = 0, t_end = 0;
    int time_iterations = 0;
    int magnetic_field_iterations = 0;
    Output_parameters output_parameters;
    Device_parameters device;
    bool keep_fields = false;  // 3D arrays held on the host (output requested)
    std::ostream &out;
    std::vector<Member> members;
    qp_wake batch = nullptr;
    double last_wake_ms_ = 0;
};
