#include "profiles.h"

#include <cmath>

profile3d gaussian3d(double amplitude, vector3d width, vector3d r0) {
    return [=](double x, double y, double z) {
        return amplitude * std::exp(-(x - r0.x) * (x - r0.x) / width.x / width.x -
                                    (y - r0.y) * (y - r0.y) / width.y / width.y -
                                    (z - r0.z) * (z - r0.z) / width.z / width.z);
    };
}

profile3d constant3d(double amplitude) {
    return [=](double, double, double) { return amplitude; };
}

profile3d parabolic3d(double amplitude, double xsize, double rsize, vector3d r0) {
    return [=](double x, double y, double z) {
        const double r2 = (y - r0.y) * (y - r0.y) + (z - r0.z) * (z - r0.z);
        const double dx2 = (x - r0.x) * (x - r0.x);
        if (r2 <= rsize * rsize && dx2 < xsize * xsize) return amplitude * (1 - dx2 / xsize / xsize) * (1 - r2 / rsize / rsize);
        return 0.0;
    };
}

profile3d cylindrical3d(double amplitude, double xsize, double rsize, vector3d r0) {
    return [=](double x, double y, double z) {
        const double r2 = (y - r0.y) * (y - r0.y) + (z - r0.z) * (z - r0.z);
        const double dx2 = (x - r0.x) * (x - r0.x);
        return (r2 <= rsize * rsize && dx2 < xsize * xsize) ? amplitude : 0.0;
    };
}

profile2d constant2d(double amplitude) {
    return [=](double, double) { return amplitude; };
}

profile2d powerlaw2d(double power, double radius, double x0, double y0, double min_factor, double max_factor,
                     double amplitude) {
    return [=](double x, double y) {
        const double r = std::sqrt((x - x0) * (x - x0) + (y - y0) * (y - y0));
        double v = std::pow(r / radius, power);
        if (v > max_factor) v = max_factor;
        if (v < min_factor) v = min_factor;
        return amplitude * v;
    };
}

profile2d channel2d(double base, double curvature, double x0, double y0, double amplitude) {
    return [=](double x, double y) {
        const double r2 = (x - x0) * (x - x0) + (y - y0) * (y - y0);
        return amplitude * (base + curvature * r2);
    };
}

profile2d ramp2d(double slope, double y0, double amplitude) {
    return [=](double y, double) {
        const double v = 1.0 + slope * (y - y0);
        return amplitude * (v > 0.0 ? v : 0.0);
    };
}
