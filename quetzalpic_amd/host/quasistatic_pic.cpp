// quasistatic_pic -- command-line driver (src/quasistatic_pic.cpp:8-47):
// one TOML input deck, runs the simulation, prints the wall time.  A deck with
// [[wake]] members runs them as one batch (Wake_batch).
#include <chrono>
#include <iostream>

#include "Config_reader.h"
#include "System_3d.h"
#include "Wake_batch.h"

int main(int argc, char const *argv[]) {
    if (argc < 2) {
        std::cout << "Provide a path to the input .toml file as the first argument. Aborting." << std::endl;
        return -1;
    }
    if (argc > 2)
        std::cout << "WARNING: more than one command-line argument are not supported. Additional arguments will be ignored"
                  << std::endl;
    std::cout << "----------------------------------------" << std::endl;
    try {
        System_parameters params = Config_reader(argv[1], std::cout).get_parameters();
        std::cout << "----------------------------------------" << std::endl;
        std::cout << "Device: " << qp_version() << std::endl;
        std::cout << "----------------------------------------" << std::endl;
        auto t0 = std::chrono::steady_clock::now();
        if (params.wakes.empty()) {
            System_3d system{params, std::cout};
            system.run();
        } else {  // [[wake]]: one batch of independent wakes
            Wake_batch batch{params, std::cout};
            batch.run();
        }
        auto t1 = std::chrono::steady_clock::now();
        std::cout << "----------------------------------------" << std::endl;
        std::cout << "Finished in " << std::chrono::duration_cast<std::chrono::seconds>(t1 - t0).count() << " s"
                  << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
