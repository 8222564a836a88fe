// Output_writer.h -- HDF5 output with the reference's file layout
// (src/Output_writer.h:11-26, src/Output_writer.cpp:7-251): Fields_NNN.h5
// (3D datasets, per-slice hyperslabs), Fields_xy_NNN.h5 (z0-plane
// interpolations), Bunch_NNN.h5 / Photons_NNN.h5 (per-particle 1D datasets);
// attributes "steps", "origin" and, for planes, "plane", "plane_coordinate".
// Written with the HDF5 C API.
#pragma once

#include <string>
#include <vector>

#include "System_parameters.h"
#include "arrays.h"

class Output_writer {
public:
    // Files <prefix>Fields_NNN.h5 etc. in the working directory (prefix empty
    // for the reference's names; a [[wake]] member's name + "_" in a batch).
    Output_writer(Output_parameters output_parameters, int count, const std::string &prefix = "");
    ~Output_writer();
    Output_writer(const Output_writer &) = delete;
    Output_writer &operator=(const Output_writer &) = delete;
    Output_writer(Output_writer &&o) noexcept;

    void initialize_slice_array(ivector3d size, vector3d steps, const array2d &array, const std::string &name);
    void write_array(const array3d &array, const std::string &name);
    void write_slice(const array2d &array, const std::string &name, int slice_index);
    void write_bunch_parameters(const std::vector<bunch_particle_3d> &particles, double weight_norm);
    void write_photon_parameters(const std::vector<bunch_particle_3d> &photons, double weight_norm);

private:
    Output_parameters output_parameters;
    long long fields_file = -1, fields_xy_file = -1, bunch_file = -1, photons_file = -1;
};

// z0-plane interpolation of a 3D array / y-line of a YZ plane
// (src/array_utils.cpp:7-45).  An index k0+1 past the last node uses k0.
array2d calculate_xy_slice(const array3d &array, double z0);
array1d calculate_y_slice(const array2d &array, double z0);

// Read-back helpers (tests and post-processing without h5py), C linkage so
// they can be called through ctypes.  Return 0 on success, -1 on failure.
extern "C" {
// rank and dims (up to 4) of a double dataset
int qpic_h5_dataset_shape(const char *path, const char *dataset, int *rank, unsigned long long *dims);
int qpic_h5_read_dataset(const char *path, const char *dataset, double *out, unsigned long long count);
// double-valued attribute (count values) or int-valued attribute (count = 1)
int qpic_h5_read_attr(const char *path, const char *dataset, const char *attr, double *out, unsigned long long count);
int qpic_h5_read_attr_int(const char *path, const char *dataset, const char *attr, int *out);
}
