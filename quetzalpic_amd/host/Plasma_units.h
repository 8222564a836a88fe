// Plasma_units.h -- CGS constants and the plasma-unit conversions the reference
// uses for diagnostics (src/constants.h:3-12, src/Plasma_units.h:7-31).
#pragma once

#include <cmath>

constexpr double QP_PI = 3.141592653589793238;
constexpr double ELECTRON_MASS_CGS = 9.1093837015e-28;
constexpr double PLANCK_CONST_CGS = 6.62607015e-27;
constexpr double PLANCK_CONST_BAR_CGS = PLANCK_CONST_CGS / (2 * QP_PI);
constexpr double SPEED_OF_LIGHT_CGS = 29979245800.0;
constexpr double ELEMENTARY_CHARGE_CGS = 4.803204672997e-10;

inline double frequency_to_density_CGS(double f) {
    return ELECTRON_MASS_CGS * f * f / (4 * QP_PI * ELEMENTARY_CHARGE_CGS * ELEMENTARY_CHARGE_CGS);
}
inline double density_to_frequency(double n) {
    return std::sqrt(4 * QP_PI * ELEMENTARY_CHARGE_CGS * ELEMENTARY_CHARGE_CGS * n / ELECTRON_MASS_CGS);
}

struct Plasma_units {
    double frequency, plasma_wavenumber, wavelength, density, number_density_norm, field_schwinger;
    explicit Plasma_units(double base_frequency_SI = 1.0)
        : frequency(base_frequency_SI), plasma_wavenumber(frequency / SPEED_OF_LIGHT_CGS),
          wavelength(2 * QP_PI / plasma_wavenumber), density(frequency_to_density_CGS(frequency)),
          number_density_norm(density / plasma_wavenumber / plasma_wavenumber / plasma_wavenumber),
          field_schwinger(ELECTRON_MASS_CGS * SPEED_OF_LIGHT_CGS * SPEED_OF_LIGHT_CGS / PLANCK_CONST_BAR_CGS /
                          frequency) {}
};
