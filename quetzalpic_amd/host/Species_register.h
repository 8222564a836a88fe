// Species_register.h -- particle species by name (interface of the
// reference's src/Species_register.h:6-24).
#pragma once

#include <string>
#include <vector>

struct Species {
    double mass;
    double charge;
    double charge_to_mass_ratio;
    bool is_photon;
    std::string name;
    Species(std::string name, double charge, double mass, bool is_photon = false)
        : mass(mass), charge(charge), charge_to_mass_ratio(mass == 0.0 ? 0.0 : charge / mass),
          is_photon(is_photon), name(std::move(name)) {}
};

class Species_register {
public:
    // throws std::invalid_argument if the name is taken; returns the new id
    int register_species(std::string name, double charge, double mass, bool is_photon = false);
    bool is_registered(const std::string &name) const;
    // id of the species or -1
    int find_species(const std::string &name) const;
    const Species &operator[](int i) const { return species_[i]; }
    const std::vector<Species> &get_species() const { return species_; }

private:
    std::vector<Species> species_;
};
