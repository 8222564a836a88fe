// arrays.h -- row-major host arrays with step/origin metadata (the metadata is
// what the HDF5 writer stores as attributes; src/array1d.h, src/array2d.h,
// src/array3d.h describe the same fields).
#pragma once

#include <cstddef>
#include <vector>

#include "containers.h"

template <class T>
class array1d_t {
public:
    array1d_t() = default;
    array1d_t(int n, double step, double origin, Axis axis, vector2d axis_coordinate)
        : n_(n), data_(n), step_(step), axis_(axis) {
        if (axis == Axis::NONE || axis == Axis::X) origin_ = {origin, axis_coordinate[0], axis_coordinate[1]};
        else if (axis == Axis::Y) origin_ = {axis_coordinate[0], origin, axis_coordinate[1]};
        else origin_ = {axis_coordinate[0], axis_coordinate[1], origin};
    }
    T &operator()(int i) { return data_[i]; }
    const T &operator()(int i) const { return data_[i]; }
    int get_size() const { return n_; }
    double get_step() const { return step_; }
    vector3d get_origin_3d() const { return origin_; }
    T *data() { return data_.data(); }
    const T *data() const { return data_.data(); }

private:
    int n_ = 0;
    std::vector<T> data_;
    double step_ = 0;
    vector3d origin_{0, 0, 0};
    Axis axis_ = Axis::NONE;
};

template <class T>
class array2d_t {
public:
    array2d_t() = default;
    // 2D steps/origin, optional plane and the coordinate of that plane
    array2d_t(ivector2d n, vector2d steps = {0, 0}, vector2d origin = {0, 0}, Plane plane = Plane::NONE,
              double plane_coordinate = 0)
        : n_(n), data_((size_t)n[0] * n[1]), steps_(steps), origin_2d_(origin), plane_(plane) {
        if (plane == Plane::NONE || plane == Plane::XY) origin_ = {origin[0], origin[1], plane_coordinate};
        else if (plane == Plane::XZ) origin_ = {origin[0], plane_coordinate, origin[1]};
        else origin_ = {plane_coordinate, origin[0], origin[1]};
    }
    // 3D steps/origin projected onto the plane
    array2d_t(ivector2d n, vector3d steps, vector3d origin, Plane plane)
        : n_(n), data_((size_t)n[0] * n[1]), origin_(origin), plane_(plane) {
        if (plane == Plane::NONE || plane == Plane::XY) {
            steps_ = {steps.x, steps.y};
            origin_2d_ = {origin.x, origin.y};
        } else if (plane == Plane::XZ) {
            steps_ = {steps.x, steps.z};
            origin_2d_ = {origin.x, origin.z};
        } else {
            steps_ = {steps.y, steps.z};
            origin_2d_ = {origin.y, origin.z};
        }
    }
    T &operator()(int i, int j) { return data_[(size_t)n_[1] * i + j]; }
    const T &operator()(int i, int j) const { return data_[(size_t)n_[1] * i + j]; }
    int get_n1() const { return n_[0]; }
    int get_n2() const { return n_[1]; }
    ivector2d get_dimensions() const { return n_; }
    vector3d get_origin_3d() const { return origin_; }
    vector2d get_steps() const { return steps_; }
    vector2d get_origin_2d() const { return origin_2d_; }
    Plane get_plane() const { return plane_; }
    double get_plane_coordinate() const {
        if (plane_ == Plane::NONE || plane_ == Plane::XY) return origin_.z;
        if (plane_ == Plane::XZ) return origin_.y;
        return origin_.x;
    }
    T *data() { return data_.data(); }
    const T *data() const { return data_.data(); }

private:
    ivector2d n_{0, 0};
    std::vector<T> data_;
    vector2d steps_{0, 0};
    vector3d origin_{0, 0, 0};
    vector2d origin_2d_{0, 0};
    Plane plane_ = Plane::NONE;
};

template <class T>
class array3d_t {
public:
    array3d_t() = default;
    array3d_t(ivector3d n, vector3d d = {0, 0, 0}, vector3d origin = {0, 0, 0})
        : n_(n), data_((size_t)n.x * n.y * n.z), d_(d), origin_(origin) {}
    T &operator()(size_t i, size_t j, size_t k) { return data_[((size_t)n_.y * i + j) * n_.z + k]; }
    const T &operator()(size_t i, size_t j, size_t k) const { return data_[((size_t)n_.y * i + j) * n_.z + k]; }
    ivector3d get_dimensions() const { return n_; }
    int get_n1() const { return n_.x; }
    int get_n2() const { return n_.y; }
    int get_n3() const { return n_.z; }
    vector3d get_origin() const { return origin_; }
    vector3d get_steps() const { return d_; }
    T *slice(int i) { return data_.data() + (size_t)i * n_.y * n_.z; }
    T *data() { return data_.data(); }
    const T *data() const { return data_.data(); }
    bool empty() const { return data_.empty(); }

private:
    ivector3d n_{0, 0, 0};
    std::vector<T> data_;
    vector3d d_{0, 0, 0};
    vector3d origin_{0, 0, 0};
};

using array1d = array1d_t<double>;
using array2d = array2d_t<double>;
using array3d = array3d_t<double>;
