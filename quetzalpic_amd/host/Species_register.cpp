#include "Species_register.h"

#include <stdexcept>

int Species_register::register_species(std::string name, double charge, double mass, bool is_photon) {
    if (is_registered(name)) throw std::invalid_argument("Species [" + name + "] already registered");
    species_.emplace_back(std::move(name), charge, mass, is_photon);
    return (int)species_.size() - 1;
}

bool Species_register::is_registered(const std::string &name) const { return find_species(name) >= 0; }

int Species_register::find_species(const std::string &name) const {
    for (size_t i = 0; i < species_.size(); ++i)
        if (species_[i].name == name) return (int)i;
    return -1;
}
