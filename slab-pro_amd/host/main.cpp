// main.cpp -- the slab-pro command: one JSON file in, simulation folder out
// (src/main.cpp:29-58).  Extra options, after the JSON path:
//   --fp32 / --fp64   arithmetic of the GPU step (default: the JSON's
//                     numerical_params.precision, else FP64)
//   --timing          per-phase kernel times after the run
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

#include "slab_frontend.h"

static int fail(const char* stage) {
    std::cerr << "Error: " << stage << ": " << slab_last_error() << std::endl;
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " <json_file> [--fp32|--fp64] [--timing]" << std::endl;
        return 1;
    }
    int precision = -1, timing = 0;
    for (int k = 2; k < argc; ++k) {
        if (!std::strcmp(argv[k], "--fp32"))
            precision = SLAB_FP32;
        else if (!std::strcmp(argv[k], "--fp64"))
            precision = SLAB_FP64;
        else if (!std::strcmp(argv[k], "--timing"))
            timing = 1;
        else {
            std::cerr << "Unknown option " << argv[k] << std::endl;
            return 1;
        }
    }
    slab_sim* s = nullptr;
    if (slab_sim_load_json(argv[1], &s) != SLAB_OK) return fail("loadFromJSON");
    slab_sim_set_verbose(s, 1, timing);
    if (precision >= 0) slab_sim_set_precision(s, precision);
    int rc = 0;
    if (slab_sim_initialize(s) != SLAB_OK) {
        rc = fail("initialize");
    } else {
        std::printf("Initialization complete.\n");
        if (slab_sim_initialize_terrain(s) != SLAB_OK) {
            rc = fail("initializeTerrain");
        } else {
            std::printf("Terrain initialized.\n");
            if (slab_sim_initialize_particles(s) != SLAB_OK) {
                rc = fail("initializeParticles");
            } else {
                std::printf("Particles initialized.\n");
                std::fflush(stdout);
                if (slab_sim_simulate(s, nullptr) != SLAB_OK) rc = fail("simulate");
            }
        }
    }
    slab_sim_destroy(s);
    return rc;
}
