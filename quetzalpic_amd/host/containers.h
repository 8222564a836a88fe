// containers.h -- small value types of the host layer (same shapes as the
// reference's src/containers.h:6-60, so calling code compiles unchanged).
#pragma once

#include <array>
#include <vector>

template <class T>
struct vector3d_t {
    T x, y, z;
    vector3d_t<T> &operator/=(T v) {
        x /= v;
        y /= v;
        z /= v;
        return *this;
    }
};
using vector3d = vector3d_t<double>;
using ivector3d = vector3d_t<int>;
using vector2d = std::array<double, 2>;
using ivector2d = std::array<int, 2>;

enum class Plane { NONE = 0, XY = 1, XZ = 2, YZ = 3 };
enum class Axis { NONE = 0, X = 1, Y = 2, Z = 3 };

// Plasma macro-particle of the wake (src/containers.h:31-43).  On the device
// the library keeps its own layout; this struct is the host-side view.
struct wake_particle_2d {
    double y = 0, y_middle = 0, z = 0, z_middle = 0;
    double py = 0, pz = 0, py_next = 0, pz_next = 0;
    double n = 0, gamma = 1, px = 0;
};

// Driver-bunch particle (src/containers.h:45-60).
struct bunch_particle_3d {
    double x = 0, y = 0, z = 0;
    double px = 0, py = 0, pz = 0;
    double gamma = 1, n = 0;
    double ex = 0, ey = 0, ez = 0, by = 0, bz = 0, chi = 0;
    int species_id = 0;
};
