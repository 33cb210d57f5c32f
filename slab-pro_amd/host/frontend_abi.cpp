// frontend_abi.cpp -- the C ABI of include/slab_frontend.h over slabhost::Simulation.
#include <algorithm>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

#include "io.hpp"
#include "poisson.hpp"
#include "simulation.hpp"
#include "slab_error.hpp"
#include "slab_frontend.h"

struct slab_sim {
    slabhost::Simulation sim;
    bool terrain_ready = false, particles_ready = false, initialized = false;
};

namespace {

template <typename F>
slab_status guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        return slab::set_error(SLAB_EINVAL, e.what());
    }
}

std::vector<std::pair<double, double>> ring_of(const double* x, const double* y, int32_t n) {
    std::vector<std::pair<double, double>> r;
    r.reserve(std::max(0, n));
    for (int32_t k = 0; k < n; ++k) r.emplace_back(x[k], y[k]);
    return r;
}

} // namespace

extern "C" {

slab_status slab_sim_load_json(const char* json_path, slab_sim** out) {
    if (!json_path || !out) return slab::set_error(SLAB_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto* s = new slab_sim();
        try {
            s->sim.loadFromJSON(json_path);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return SLAB_OK;
    });
}

slab_status slab_sim_initialize(slab_sim* s) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    return guarded([&] {
        s->sim.initialize();
        s->initialized = true;
        return SLAB_OK;
    });
}

slab_status slab_sim_initialize_terrain(slab_sim* s) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    return guarded([&] {
        s->sim.initializeTerrain();
        s->terrain_ready = true;
        return SLAB_OK;
    });
}

slab_status slab_sim_initialize_particles(slab_sim* s) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (!s->terrain_ready) return slab::set_error(SLAB_ESTATE, "call slab_sim_initialize_terrain first");
    return guarded([&] {
        s->sim.initializeParticles();
        s->particles_ready = true;
        return SLAB_OK;
    });
}

slab_status slab_sim_simulate(slab_sim* s, int64_t* steps) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (!s->initialized || !s->particles_ready)
        return slab::set_error(SLAB_ESTATE, "call slab_sim_initialize and slab_sim_initialize_particles first");
    try {
        const int64_t n = s->sim.simulate();
        if (steps) *steps = n;
        return SLAB_OK;
    } catch (const std::exception& e) {
        // a failing C ABI call already left its message; keep the prefix too
        return slab::set_error(SLAB_EDEVICE, e.what());
    }
}

slab_status slab_sim_set_verbose(slab_sim* s, int32_t verbose, int32_t timing) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    s->sim.verbose = verbose != 0;
    s->sim.timing = timing != 0;
    return SLAB_OK;
}

slab_status slab_sim_set_precision(slab_sim* s, int32_t precision) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (precision != SLAB_FP64 && precision != SLAB_FP32) return slab::set_error(SLAB_EINVAL, "unknown precision");
    s->sim.cfg.precision = precision;
    return SLAB_OK;
}

slab_status slab_sim_terrain(const slab_sim* s, int32_t* nx, int32_t* ny, double* dx, double* xll, double* yll,
                             double* z) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (!s->terrain_ready) return slab::set_error(SLAB_ESTATE, "terrain not initialized");
    const auto& t = s->sim.terrain;
    if (nx) *nx = t.nx;
    if (ny) *ny = t.ny;
    if (dx) *dx = t.dx;
    if (xll) *xll = t.xll;
    if (yll) *yll = t.yll;
    if (z) std::memcpy(z, t.z.data(), t.z.size() * sizeof(double));
    return SLAB_OK;
}

int64_t slab_sim_num_particles(const slab_sim* s) {
    return (s && s->particles_ready) ? (int64_t)s->sim.particles.x.size() : -1;
}

slab_status slab_sim_particles(const slab_sim* s, double* x, double* y, double* h, double* A, double* m) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (!s->particles_ready) return slab::set_error(SLAB_ESTATE, "particles not initialized");
    const auto& p = s->sim.particles;
    const size_t n = p.x.size();
    if (x) std::memcpy(x, p.x.data(), n * sizeof(double));
    if (y) std::memcpy(y, p.y.data(), n * sizeof(double));
    if (h) std::memcpy(h, p.h.data(), n * sizeof(double));
    if (A) std::memcpy(A, p.A.data(), n * sizeof(double));
    if (m) std::memcpy(m, p.m.data(), n * sizeof(double));
    return SLAB_OK;
}

slab_status slab_sim_release(const slab_sim* s, double* area, double* volume) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (!s->particles_ready) return slab::set_error(SLAB_ESTATE, "particles not initialized");
    if (area) *area = s->sim.particles.release_area;
    if (volume) *volume = s->sim.particles.release_volume;
    return SLAB_OK;
}

slab_status slab_sim_output(const slab_sim* s, char* sim_dir, int32_t cap, int32_t* n_files, int32_t* padding) {
    if (!s) return slab::set_error(SLAB_EINVAL, "null simulation");
    if (!s->initialized) return slab::set_error(SLAB_ESTATE, "call slab_sim_initialize first");
    if (sim_dir && cap > 0) {
        const std::string& d = s->sim.sim_dir;
        const size_t n = std::min((size_t)cap - 1, d.size());
        std::memcpy(sim_dir, d.data(), n);
        sim_dir[n] = 0;
    }
    if (n_files) *n_files = s->sim.nFiles;
    if (padding) *padding = s->sim.padding;
    return SLAB_OK;
}

void slab_sim_destroy(slab_sim* s) { delete s; }

int64_t slab_poisson_disk(double dx, double Lx, double Ly, int32_t ppc, double* x, double* y, int64_t cap) {
    if (!(dx > 0) || !(Lx >= 0) || !(Ly >= 0) || ppc < 1) {
        slab::set_error(SLAB_EINVAL, "poisson_disk: bad arguments");
        return -1;
    }
    try {
        const auto s = slabhost::poisson_disk(dx, Lx, Ly, ppc);
        const int64_t n = (int64_t)s.size();
        for (int64_t k = 0; k < std::min(n, cap); ++k) {
            if (x) x[k] = s[k].first;
            if (y) y[k] = s[k].second;
        }
        return n;
    } catch (const std::exception& e) {
        slab::set_error(SLAB_ENOMEM, e.what());
        return -1;
    }
}

double slab_polygon_area(const double* x, const double* y, int32_t n) {
    if (!x || !y || n < 3) return 0.0;
    return slabhost::polygon_area(ring_of(x, y, n));
}

int32_t slab_point_in_polygon(double px, double py, const double* x, const double* y, int32_t n) {
    if (!x || !y || n < 1) return 0;
    return slabhost::point_in_polygon(px, py, ring_of(x, y, n)) ? 1 : 0;
}

slab_status slab_read_shapefile_polygon(const char* path, double* x, double* y, int32_t cap, int32_t* n) {
    if (!path || !n) return slab::set_error(SLAB_EINVAL, "null argument");
    return guarded([&] {
        const auto r = slabhost::read_shapefile_polygon(path);
        *n = (int32_t)r.size();
        for (int32_t k = 0; k < std::min(*n, cap); ++k) {
            if (x) x[k] = r[k].first;
            if (y) y[k] = r[k].second;
        }
        return SLAB_OK;
    });
}

slab_status slab_write_ply(const char* path, int64_t np, const double* x, const double* y, const double* vx,
                           const double* vy, const double* vz, const double* z, const double* h, const double* A) {
    if (!path || np < 0 || (np > 0 && (!x || !y || !vx || !vy || !vz || !z || !h || !A)))
        return slab::set_error(SLAB_EINVAL, "null argument");
    return guarded([&] {
        slabhost::write_ply(path, np, x, y, vx, vy, vz, z, h, A);
        return SLAB_OK;
    });
}

slab_status slab_write_geotiff(const char* path, int32_t nx, int32_t ny, double x0, double ytop, double dx,
                               int32_t nbands, const char* const* names, const float* const* bands, float nodata) {
    if (!path || !names || !bands || nx < 1 || ny < 1 || nbands < 1) return slab::set_error(SLAB_EINVAL, "bad argument");
    return guarded([&] {
        std::vector<std::string> nm;
        std::vector<const float*> b;
        for (int32_t k = 0; k < nbands; ++k) {
            if (!names[k] || !bands[k]) throw std::runtime_error("null band");
            nm.emplace_back(names[k]);
            b.push_back(bands[k]);
        }
        slabhost::write_geotiff(path, nx, ny, x0, ytop, dx, nm, b, nodata);
        return SLAB_OK;
    });
}

} // extern "C"
