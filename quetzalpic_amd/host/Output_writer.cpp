#include "Output_writer.h"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace {

hid_t create_file(const std::string &name) {
    hid_t f = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) throw std::runtime_error("cannot create " + name);
    return f;
}

std::string numbered(const char *stem, int count) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%s_%03d.h5", stem, count);
    return buf;
}

void write_attr_doubles(hid_t obj, const char *name, const double *v, hsize_t n) {
    hid_t sp = H5Screate_simple(1, &n, nullptr);
    hid_t a = H5Acreate2(obj, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_DOUBLE, v);
    H5Aclose(a);
    H5Sclose(sp);
}

void write_plane_attrs(hid_t obj, Plane plane, double coord) {
    if (plane == Plane::NONE) return;
    hid_t sp = H5Screate(H5S_SCALAR);
    int id = (int)plane;
    hid_t a = H5Acreate2(obj, "plane", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_INT, &id);
    H5Aclose(a);
    a = H5Acreate2(obj, "plane_coordinate", H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_DOUBLE, &coord);
    H5Aclose(a);
    H5Sclose(sp);
}

hid_t create_dataset(hid_t file, const std::string &name, int rank, const hsize_t *dims) {
    hid_t sp = H5Screate_simple(rank, dims, nullptr);
    hid_t d = H5Dcreate2(file, name.c_str(), H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(sp);
    if (d < 0) throw std::runtime_error("cannot create dataset " + name);
    return d;
}

// write_array(array3d) of src/Output_writer.cpp:7-26
void write_array3d(hid_t file, const array3d &a, const std::string &name) {
    const hsize_t dims[3] = {(hsize_t)a.get_n1(), (hsize_t)a.get_n2(), (hsize_t)a.get_n3()};
    hid_t d = create_dataset(file, name, 3, dims);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.data());
    const vector3d s = a.get_steps(), o = a.get_origin();
    const double sv[3] = {s.x, s.y, s.z}, ov[3] = {o.x, o.y, o.z};
    write_attr_doubles(d, "steps", sv, 3);
    write_attr_doubles(d, "origin", ov, 3);
    H5Dclose(d);
}

// write_array(array2d) of src/Output_writer.cpp:28-60
void write_array2d(hid_t file, const array2d &a, const std::string &name) {
    const hsize_t dims[2] = {(hsize_t)a.get_n1(), (hsize_t)a.get_n2()};
    hid_t d = create_dataset(file, name, 2, dims);
    H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.data());
    const vector2d s = a.get_steps(), o = a.get_origin_2d();
    write_attr_doubles(d, "steps", s.data(), 2);
    write_attr_doubles(d, "origin", o.data(), 2);
    write_plane_attrs(d, a.get_plane(), a.get_plane_coordinate());
    H5Dclose(d);
}

// Hyperslab write of one slice: a 2D plane into [slice, :, :] of a 3D dataset
// (rank 3) or a line into [slice, :] of a 2D dataset (rank 2).
void write_slice_rows(hid_t file, const std::string &name, hsize_t slice, const double *data, hsize_t n1, hsize_t n2,
                      int rank) {
    hid_t d = H5Dopen2(file, name.c_str(), H5P_DEFAULT);
    if (d < 0) throw std::runtime_error("no dataset " + name);
    hid_t fs = H5Dget_space(d);
    hid_t ms;
    if (rank == 3) {
        const hsize_t start[3] = {slice, 0, 0}, count[3] = {1, n1, n2};
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
        const hsize_t m[2] = {n1, n2};
        ms = H5Screate_simple(2, m, nullptr);
    } else {
        const hsize_t start[2] = {slice, 0}, count[2] = {1, n1};
        H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, nullptr, count, nullptr);
        ms = H5Screate_simple(1, &n1, nullptr);
    }
    H5Dwrite(d, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, data);
    H5Sclose(ms);
    H5Sclose(fs);
    H5Dclose(d);
}

void write_particles(hid_t file, const std::vector<bunch_particle_3d> &p, double weight_norm) {
    const hsize_t n = p.size();
    std::vector<double> v(n);
    auto put = [&](const char *name, double (*get)(const bunch_particle_3d &)) {
        for (hsize_t i = 0; i < n; ++i) v[i] = get(p[i]);
        hid_t d = create_dataset(file, name, 1, &n);
        if (n) H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
        H5Dclose(d);
    };
    put("x", [](const bunch_particle_3d &q) { return q.x; });
    put("y", [](const bunch_particle_3d &q) { return q.y; });
    put("z", [](const bunch_particle_3d &q) { return q.z; });
    put("px", [](const bunch_particle_3d &q) { return q.px; });
    put("py", [](const bunch_particle_3d &q) { return q.py; });
    put("pz", [](const bunch_particle_3d &q) { return q.pz; });
    put("gamma", [](const bunch_particle_3d &q) { return q.gamma; });
    put("ex", [](const bunch_particle_3d &q) { return q.ex; });
    put("ey", [](const bunch_particle_3d &q) { return q.ey; });
    put("ez", [](const bunch_particle_3d &q) { return q.ez; });
    put("by", [](const bunch_particle_3d &q) { return q.by; });
    put("bz", [](const bunch_particle_3d &q) { return q.bz; });
    put("chi", [](const bunch_particle_3d &q) { return q.chi; });
    for (hsize_t i = 0; i < n; ++i) v[i] = p[i].n * weight_norm;
    hid_t d = create_dataset(file, "weight", 1, &n);
    if (n) H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
    H5Dclose(d);
}

}  // namespace

array2d calculate_xy_slice(const array3d &a, double z0) {
    const ivector3d n = a.get_dimensions();
    const vector3d o = a.get_origin(), s = a.get_steps();
    array2d r({n.x, n.y}, vector2d{s.x, s.y}, vector2d{o.x, o.y}, Plane::XY, z0);
    int k0 = (int)std::floor((z0 - o.z) / s.z);
    const double zf = (z0 - o.z) / s.z - k0;
    if (k0 < 0) k0 = 0;
    if (k0 > n.z - 1) k0 = n.z - 1;
    const int k1 = k0 + 1 < n.z ? k0 + 1 : k0;
    for (int i = 0; i < n.x; ++i)
        for (int j = 0; j < n.y; ++j) r(i, j) = (1 - zf) * a(i, j, k0) + zf * a(i, j, k1);
    return r;
}

array1d calculate_y_slice(const array2d &a, double z0) {
    if (a.get_plane() != Plane::YZ) throw std::invalid_argument("Plane should be YZ");
    const ivector2d n = a.get_dimensions();
    const vector3d o = a.get_origin_3d();
    const vector2d s = a.get_steps();
    array1d r(n[0], s[0], o.y, Axis::Y, vector2d{o.x, z0});
    int k0 = (int)std::floor((z0 - o.z) / s[1]);
    const double zf = (z0 - o.z) / s[1] - k0;
    if (k0 < 0) k0 = 0;
    if (k0 > n[1] - 1) k0 = n[1] - 1;
    const int k1 = k0 + 1 < n[1] ? k0 + 1 : k0;
    for (int j = 0; j < n[0]; ++j) r(j) = (1 - zf) * a(j, k0) + zf * a(j, k1);
    return r;
}

Output_writer::Output_writer(Output_parameters p, int count, const std::string &prefix) : output_parameters(p) {
    if (p.output3d) fields_file = create_file(prefix + numbered("Fields", count));
    if (p.output_xy) fields_xy_file = create_file(prefix + numbered("Fields_xy", count));
    if (p.output_bunch) bunch_file = create_file(prefix + numbered("Bunch", count));
    if (p.output_photons) photons_file = create_file(prefix + numbered("Photons", count));
}

Output_writer::Output_writer(Output_writer &&o) noexcept
    : output_parameters(o.output_parameters), fields_file(o.fields_file), fields_xy_file(o.fields_xy_file),
      bunch_file(o.bunch_file), photons_file(o.photons_file) {
    o.fields_file = o.fields_xy_file = o.bunch_file = o.photons_file = -1;
}

Output_writer::~Output_writer() {
    for (long long f : {fields_file, fields_xy_file, bunch_file, photons_file})
        if (f >= 0) H5Fclose((hid_t)f);
}

// src/Output_writer.cpp:62-109, 185-197
void Output_writer::initialize_slice_array(ivector3d size, vector3d steps, const array2d &array,
                                           const std::string &name) {
    const vector3d o = array.get_origin_3d();
    if (output_parameters.output3d) {
        const hsize_t dims[3] = {(hsize_t)size.x, (hsize_t)size.y, (hsize_t)size.z};
        hid_t d = create_dataset((hid_t)fields_file, name, 3, dims);
        const double sv[3] = {steps.x, steps.y, steps.z}, ov[3] = {o.x, o.y, o.z};
        write_attr_doubles(d, "steps", sv, 3);
        write_attr_doubles(d, "origin", ov, 3);
        H5Dclose(d);
    }
    if (output_parameters.output_xy) {
        const hsize_t dims[2] = {(hsize_t)size.x, (hsize_t)size.y};
        hid_t d = create_dataset((hid_t)fields_xy_file, name, 2, dims);
        const double sv[2] = {steps.x, steps.y}, ov[2] = {o.x, o.y};
        write_attr_doubles(d, "steps", sv, 2);
        write_attr_doubles(d, "origin", ov, 2);
        write_plane_attrs(d, Plane::XY, output_parameters.z0);
        H5Dclose(d);
    }
}

void Output_writer::write_array(const array3d &array, const std::string &name) {
    if (output_parameters.output3d) write_array3d((hid_t)fields_file, array, name);
    if (output_parameters.output_xy) write_array2d((hid_t)fields_xy_file, calculate_xy_slice(array, output_parameters.z0), name);
}

void Output_writer::write_slice(const array2d &array, const std::string &name, int slice_index) {
    if (output_parameters.output3d)
        write_slice_rows((hid_t)fields_file, name, (hsize_t)slice_index, array.data(), (hsize_t)array.get_n1(),
                         (hsize_t)array.get_n2(), 3);
    if (output_parameters.output_xy) {
        const array1d line = calculate_y_slice(array, output_parameters.z0);
        write_slice_rows((hid_t)fields_xy_file, name, (hsize_t)slice_index, line.data(), (hsize_t)line.get_size(), 1,
                         2);
    }
}

void Output_writer::write_bunch_parameters(const std::vector<bunch_particle_3d> &p, double weight_norm) {
    if (output_parameters.output_bunch) write_particles((hid_t)bunch_file, p, weight_norm);
}

void Output_writer::write_photon_parameters(const std::vector<bunch_particle_3d> &p, double weight_norm) {
    if (output_parameters.output_photons) write_particles((hid_t)photons_file, p, weight_norm);
}

extern "C" {

int qpic_h5_dataset_shape(const char *path, const char *dataset, int *rank, unsigned long long *dims) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        return -1;
    }
    hid_t s = H5Dget_space(d);
    hsize_t dd[8];
    const int r = H5Sget_simple_extent_dims(s, dd, nullptr);
    *rank = r;
    for (int i = 0; i < r && i < 4; ++i) dims[i] = dd[i];
    H5Sclose(s);
    H5Dclose(d);
    H5Fclose(f);
    return r < 0 || r > 4 ? -1 : 0;
}

int qpic_h5_read_dataset(const char *path, const char *dataset, double *out, unsigned long long count) {
    int rank = 0;
    unsigned long long dims[4] = {1, 1, 1, 1};
    if (qpic_h5_dataset_shape(path, dataset, &rank, dims) != 0) return -1;
    unsigned long long n = 1;
    for (int i = 0; i < rank; ++i) n *= dims[i];
    if (n != count) return -1;
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
    const herr_t e = H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out);
    H5Dclose(d);
    H5Fclose(f);
    return e < 0 ? -1 : 0;
}

static int read_attr(const char *path, const char *dataset, const char *attr, hid_t type, void *out,
                     unsigned long long count) {
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        return -1;
    }
    hid_t a = H5Aopen(d, attr, H5P_DEFAULT);
    int rc = -1;
    if (a >= 0) {
        hid_t s = H5Aget_space(a);
        const hssize_t n = H5Sget_simple_extent_npoints(s);
        if (n == (hssize_t)count && H5Aread(a, type, out) >= 0) rc = 0;
        H5Sclose(s);
        H5Aclose(a);
    }
    H5Dclose(d);
    H5Fclose(f);
    return rc;
}

int qpic_h5_read_attr(const char *path, const char *dataset, const char *attr, double *out, unsigned long long count) {
    return read_attr(path, dataset, attr, H5T_NATIVE_DOUBLE, out, count);
}

int qpic_h5_read_attr_int(const char *path, const char *dataset, const char *attr, int *out) {
    return read_attr(path, dataset, attr, H5T_NATIVE_INT, out, 1);
}

}  // extern "C"
