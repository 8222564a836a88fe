// profiles.h -- analytic bunch, laser and plasma profiles (host side).
// Same closures as src/profiles.h:1-10 plus the parabolic channel extension.
#pragma once

#include <functional>

#include "containers.h"

using profile3d = std::function<double(double, double, double)>;
using profile2d = std::function<double(double, double)>;

// amplitude * exp(-sum (r - r0)^2 / width^2) (src/profiles.cpp:4-11)
profile3d gaussian3d(double amplitude, vector3d width, vector3d r0);
profile3d constant3d(double amplitude = 1.0);
// amplitude (1 - dx^2/xsize^2)(1 - r^2/rsize^2) inside the ellipsoid (src/profiles.cpp:19-30)
profile3d parabolic3d(double amplitude, double xsize, double rsize, vector3d r0);
profile3d cylindrical3d(double amplitude, double xsize, double rsize, vector3d r0);
profile2d constant2d(double amplitude = 1.0);
// amplitude * clamp((r/radius)^power, min, max) (src/profiles.cpp:51-63)
profile2d powerlaw2d(double power, double radius, double x0, double y0, double min_factor = 0.0,
                     double max_factor = 1e10, double amplitude = 1.0);
// Extension (no reference counterpart): amplitude * (base + curvature r^2), a
// parabolic density channel.
profile2d channel2d(double base, double curvature, double x0, double y0, double amplitude = 1.0);
// Extension: amplitude * max(0, 1 + slope (y - y0)), a linear density ramp across the box.
profile2d ramp2d(double slope, double y0, double amplitude = 1.0);
