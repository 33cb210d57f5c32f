// simulation.hpp -- the configuration front end of the solver: JSON input,
// simulation folders, terrain and particle initialisation, the output-frame
// driver and the .ply / GeoTIFF writers.  It mirrors the reference's
// Simulation class (src/simulation.hpp, src/simulation.cpp:54-515,
// src/save_function.cpp) with the time step itself done on the GPU through
// the C ABI of slab_abi.h.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "io.hpp"

namespace slabhost {

enum class FrictionType { Coulomb = 0, Voellmy = 1, MuI = 2 };
enum class TerrainType { DEM = 0, Analytical = 1 };
enum class ReleaseShapeType { Rectangle = 0, Ellipse = 1, Polygon = 2, Raster = 3 };
enum class SamplingMethod { Uniform = 0, Poisson = 1 };

// Everything loadFromJSON reads (src/simulation.cpp:118-208).
struct SimConfig {
    std::string json_path, project_dir, sim_name;
    double gravity = 9.81, rho = 1000.0;
    double output_dt = 1.0, total_time = 1.0;
    double C_cfl = 0.5, dt_max = 0.01, flip_ratio = 0.95;
    SamplingMethod sampling = SamplingMethod::Uniform;
    FrictionType friction = FrictionType::Coulomb;
    double mu = 0, xi = 0, mu1 = 0, mu2 = 0, I0 = 0, d = 0, phi = 0;
    TerrainType terrain = TerrainType::DEM;
    std::string dem_path;
    double terrain_Lx = 0, terrain_Ly = 0, dx = 1.0;
    // analytical profile: 0 = the reference's 5-degree plane in x and y;
    // 1 = plane of slope theta for x <= x_t, circular arc of radius R, flat
    // runout at z_f (z and slope continuous)
    int profile = 0;
    double prof_theta_deg = 35.0, prof_x_t = 600.0, prof_R = 300.0, prof_z_f = 0.0;
    ReleaseShapeType release = ReleaseShapeType::Rectangle;
    double h0 = 0;
    double rect_cx = 0, rect_cy = 0, rect_Lx = 0, rect_Ly = 0, rect_rot = 0; // rot in radians
    double ell_cx = 0, ell_cy = 0, ell_rx = 0, ell_ry = 0, ell_rot = 0;
    std::string polygon_path, raster_path;
    double raster_h_min = 0;
    int ppc = 4;
    int precision = 0; // slab_precision: 0 FP64, 1 FP32
};

struct Terrain {
    int nx = 0, ny = 0;
    double dx = 1.0, xll = 0, yll = 0;
    std::vector<double> x, y, z; // z[i*ny + j]
};

struct ParticleSet {
    std::vector<double> x, y, h, A, m;
    double release_area = 0, release_volume = 0;
};

class Simulation {
  public:
    SimConfig cfg;
    Terrain terrain;
    ParticleSet particles;
    int nFiles = 0, padding = 0;
    std::string sim_dir, particle_dir, raster_dir;
    bool verbose = true;
    bool timing = false;

    void loadFromJSON(const std::string& path);
    void initialize();
    void initializeTerrain();
    void initializeParticles();
    // Runs the time loop on the GPU, writing particle/raster frames and the
    // max-value raster; returns the number of time steps.
    int64_t simulate();

    // Release height at a position from the release raster (0 outside it or
    // where a corner is nodata), src/complex_terrain.cpp:132-183.
    double releaseHeight(double px, double py) const;

    std::string frameName(const char* stem, int frame, const char* ext) const;

  private:
    RasterData release_raster_;
    double rr_x0_ = 0, rr_y0_ = 0;
    std::vector<std::pair<double, double>> polygon_;
};

// Polygon area of a ring (closed if it is not), in the x_i (y_{i+1} -
// y_{i-1}) form of the shoelace formula.
double polygon_area(const std::vector<std::pair<double, double>>& ring);
// Even-odd ray test with the reference's 1e-12 guard (src/polygon_tools.cpp:87-100).
bool point_in_polygon(double x, double y, const std::vector<std::pair<double, double>>& poly);
// x_0 = 0, x_k = k*dx until x_k >= length (src/complex_terrain.cpp:67-82).
std::vector<double> generate_grid(double length, double dx);

} // namespace slabhost
