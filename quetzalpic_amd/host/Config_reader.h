// Config_reader.h -- TOML input deck -> System_parameters with the reference's
// keys, defaults and errors (src/Config_reader.h:9-14, src/Config_reader.cpp:18-258).
#pragma once

#include <exception>
#include <iostream>
#include <string>

#include "System_parameters.h"
#include "toml_lite.h"

// <a^2> of the deck's Gaussian laser: a0^2/2 with the full widths at 1/e^2
// (sigma) turned into Gaussian parameters, centred at r0.
profile3d laser_a_sqr(double a0, const double sigma[3], const double r0[3]);

class Config_exception : public std::exception {
public:
    explicit Config_exception(std::string message) : message_(std::move(message)) {}
    const char *what() const noexcept override { return message_.c_str(); }

private:
    std::string message_;
};

class Config_reader {
public:
    Config_reader(const std::string &filename, std::ostream &out);
    // Parse TOML text directly (no file).
    static Config_reader from_string(const std::string &text, std::ostream &out);
    System_parameters get_parameters() { return params; }

private:
    Config_reader(toml_lite::TablePtr config, std::ostream &out);
    void read_all();
    template <class T>
    T read_value(const std::string &name, const toml_lite::TablePtr &table);
    template <class T>
    T read_value(const std::string &name, T default_value, const toml_lite::TablePtr &table);
    void init_laser();
    void parse_laser_table(const toml_lite::TablePtr &t, double &a0, double sigma[3], double r0[3]);
    profile2d parse_plasma_table(const toml_lite::TablePtr &t);
    void init_wakes();
    Bunch_parameters parse_bunch_table(const toml_lite::TablePtr &t);
    void init_bunch();
    void init_output_parameters();
    void init_plasma_profile();
    void init_device();

    System_parameters params;
    toml_lite::TablePtr config;
    std::ostream &out;
};
