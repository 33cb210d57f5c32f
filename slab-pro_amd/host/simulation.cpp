// simulation.cpp -- configuration front end (see simulation.hpp).
#include "simulation.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "json.hpp"
#include "poisson.hpp"
#include "slab_abi.h"

namespace fs = std::filesystem;

namespace slabhost {

// --- geometry helpers -------------------------------------------------------

double polygon_area(const std::vector<std::pair<double, double>>& ring_in) {
    if (ring_in.size() < 3) return 0.0;
    std::vector<std::pair<double, double>> ring = ring_in;
    if (ring.front() != ring.back()) ring.push_back(ring.front());
    const size_t n = ring.size();
    double acc = ring[0].first * (ring[1].second - ring[n - 1].second);
    for (size_t i = 1; i + 1 < n; ++i) acc += ring[i].first * (ring[i + 1].second - ring[i - 1].second);
    acc += ring[n - 1].first * (ring[0].second - ring[n - 2].second);
    return 0.5 * std::fabs(acc);
}

bool point_in_polygon(double x, double y, const std::vector<std::pair<double, double>>& poly) {
    const size_t n = poly.size();
    bool inside = false;
    for (size_t i = 0, j = n - 1; i < n; j = i++) {
        const double xi = poly[i].first, yi = poly[i].second;
        const double xj = poly[j].first, yj = poly[j].second;
        if (((yi > y) != (yj > y)) && (x < (xj - xi) * (y - yi) / ((yj - yi) + 1e-12) + xi)) inside = !inside;
    }
    return inside;
}

std::vector<double> generate_grid(double length, double dx) {
    std::vector<double> g{0.0};
    double x = 0.0;
    for (int k = 1; x < length; ++k) {
        x = (double)k * dx;
        g.push_back(x);
    }
    return g;
}

// --- configuration ----------------------------------------------------------

static FrictionType parse_friction(const std::string& s) {
    if (s == "COULOMB") return FrictionType::Coulomb;
    if (s == "VOELLMY") return FrictionType::Voellmy;
    if (s == "MU_I") return FrictionType::MuI;
    throw std::runtime_error("Unknown friction type: " + s);
}
static TerrainType parse_terrain(const std::string& s) {
    if (s == "DEM") return TerrainType::DEM;
    if (s == "ANALYTICAL") return TerrainType::Analytical;
    throw std::runtime_error("Unknown terrain type: " + s);
}
static ReleaseShapeType parse_release(const std::string& s) {
    if (s == "RECTANGLE") return ReleaseShapeType::Rectangle;
    if (s == "ELLIPSE") return ReleaseShapeType::Ellipse;
    if (s == "POLYGON") return ReleaseShapeType::Polygon;
    if (s == "RASTER") return ReleaseShapeType::Raster;
    throw std::runtime_error("Unknown release shape type: " + s);
}

void Simulation::loadFromJSON(const std::string& path) {
    if (!fs::exists(path)) throw std::runtime_error("JSON file not found: " + path);
    std::ifstream f(path);
    std::stringstream buf;
    buf << f.rdbuf();
    const Json j = Json::parse(buf.str());
    SimConfig c;
    c.json_path = path;
    c.project_dir = j.at("project_dir").str();
    c.sim_name = j.at("sim_name").str();
    const Json& phys = j.at("physical_params");
    c.gravity = phys.at("gravity").num();
    c.rho = phys.at("rho").num();
    const Json& tp = j.at("time_params");
    c.output_dt = tp.at("output_dt").num();
    c.total_time = tp.at("total_time").num();
    const Json& np = j.at("numerical_params");
    c.C_cfl = np.at("C_cfl").num();
    c.dt_max = np.at("dt_max").num();
    c.flip_ratio = np.at("flip_ratio").num();
    const std::string method = np.at("sampling_method").str();
    if (method == "UNIFORM")
        c.sampling = SamplingMethod::Uniform;
    else if (method == "POISSON")
        c.sampling = SamplingMethod::Poisson;
    else
        throw std::runtime_error("Unknown sampling method: " + method);
    // optional: arithmetic of the GPU step ("FP64" default, or "FP32")
    if (np.contains("precision")) {
        const std::string p = np.at("precision").str();
        if (p == "FP64")
            c.precision = SLAB_FP64;
        else if (p == "FP32")
            c.precision = SLAB_FP32;
        else
            throw std::runtime_error("Unknown precision: " + p);
    }
    const Json& fr = j.at("friction");
    c.friction = parse_friction(fr.at("type").str());
    if (c.friction == FrictionType::Coulomb) {
        c.mu = fr.at("coulomb").at("mu").num();
    } else if (c.friction == FrictionType::Voellmy) {
        c.mu = fr.at("voellmy").at("mu").num();
        c.xi = fr.at("voellmy").at("xi").num();
    } else {
        const Json& m = fr.at("mu_I");
        c.mu1 = m.at("mu_1").num();
        c.mu2 = m.at("mu_2").num();
        c.I0 = m.at("I0").num();
        c.d = m.at("d").num();
        c.phi = m.at("phi").num();
    }
    const Json& te = j.at("terrain");
    c.terrain = parse_terrain(te.at("type").str());
    if (c.terrain == TerrainType::DEM) {
        c.dem_path = te.at("dem").at("filepath").str();
    } else {
        const Json& a = te.at("analytical");
        c.terrain_Lx = a.at("Lx").num();
        c.terrain_Ly = a.at("Ly").num();
        c.dx = a.at("dx").num();
        // optional extension: "profile" selects another analytical terrain
        if (a.contains("profile")) {
            const std::string pr = a.at("profile").str();
            if (pr == "PLANE") {
                c.profile = 0;
            } else if (pr == "PLANE_ARC_FLAT") {
                c.profile = 1;
                if (a.contains("theta_deg")) c.prof_theta_deg = a.at("theta_deg").num();
                if (a.contains("x_t")) c.prof_x_t = a.at("x_t").num();
                if (a.contains("R")) c.prof_R = a.at("R").num();
                if (a.contains("z_f")) c.prof_z_f = a.at("z_f").num();
            } else {
                throw std::runtime_error("Unknown analytical profile: " + pr);
            }
        }
    }
    const Json& re = j.at("release");
    c.release = parse_release(re.at("type").str());
    c.h0 = re.at("h0").num();
    if (c.release == ReleaseShapeType::Rectangle) {
        const Json& r = re.at("rectangle");
        c.rect_cx = r.at("center")[0].num();
        c.rect_cy = r.at("center")[1].num();
        c.rect_Lx = r.at("Lx").num();
        c.rect_Ly = r.at("Ly").num();
        c.rect_rot = r.at("rotation_deg").num() * M_PI / 180.0;
    } else if (c.release == ReleaseShapeType::Ellipse) {
        const Json& e = re.at("ellipse");
        c.ell_cx = e.at("center")[0].num();
        c.ell_cy = e.at("center")[1].num();
        c.ell_rx = e.at("rx").num();
        c.ell_ry = e.at("ry").num();
        c.ell_rot = e.at("rotation_deg").num() * M_PI / 180.0;
    } else if (c.release == ReleaseShapeType::Polygon) {
        c.polygon_path = re.at("polygon").at("filepath").str();
    } else {
        c.raster_path = re.at("raster").at("filepath").str();
        c.raster_h_min = re.at("raster").at("h_min").num();
    }
    cfg = c;
}

std::string Simulation::frameName(const char* stem, int frame, const char* ext) const {
    std::ostringstream ss;
    ss << stem << std::setw(padding) << std::setfill('0') << frame << ext;
    return ss.str();
}

void Simulation::initialize() {
    const fs::path root = fs::path(cfg.project_dir) / cfg.sim_name;
    try {
        if (fs::create_directories(root)) {
            if (verbose) std::cout << "Directory " << root << " has been created." << std::endl;
        } else if (verbose) {
            std::cerr << "Directory " << root << " already exists or could not be created." << std::endl;
        }
        fs::create_directories(root / "rasters");
        fs::create_directories(root / "particles");
    } catch (const std::exception& e) {
        std::cerr << "Trying to create directory: " << root << std::endl;
        std::cerr << "Filesystem error: " << e.what() << std::endl;
    }
    sim_dir = root.string();
    particle_dir = (root / "particles").string();
    raster_dir = (root / "rasters").string();
    nFiles = (int)std::ceil(cfg.total_time / cfg.output_dt);
    padding = (int)std::floor(std::log10(nFiles)) + 1;
    try {
        const fs::path dest = root / fs::path(cfg.json_path).filename();
        fs::copy_file(cfg.json_path, dest, fs::copy_options::overwrite_existing);
        if (verbose) std::cout << "Copied JSON to " << dest << "\n";
    } catch (const std::exception& e) {
        std::cerr << "Failed to copy JSON file: " << e.what() << "\n";
    }
}

void Simulation::initializeTerrain() {
    Terrain t;
    if (cfg.terrain == TerrainType::DEM) {
        RasterData r = read_asc(cfg.dem_path);
        t.nx = r.ncols;
        t.ny = r.nrows;
        t.x = std::move(r.x);
        t.y = std::move(r.y);
        t.z = std::move(r.z);
        t.xll = r.xllcorner;
        t.yll = r.yllcorner;
        cfg.dx = r.cellsize;
    } else {
        t.x = generate_grid(cfg.terrain_Lx, cfg.dx);
        t.y = generate_grid(cfg.terrain_Ly, cfg.dx);
        t.nx = (int)t.x.size();
        t.ny = (int)t.y.size();
        t.z.resize((size_t)t.nx * t.ny);
        if (cfg.profile == 0) {
            // the reference's analytical terrain: a plane falling 5 degrees
            // along x and along y (src/simulation.cpp:181-185)
            const double th = 5.0 * M_PI / 180.0;
            for (int i = 0; i < t.nx; ++i)
                for (int j = 0; j < t.ny; ++j)
                    t.z[(size_t)i * t.ny + j] = -t.x[i] * std::tan(th) - t.y[j] * std::tan(th);
        } else {
            // plane / circular arc / flat runout along x, constant along y
            const double th = cfg.prof_theta_deg * M_PI / 180.0, R = cfg.prof_R, xt = cfg.prof_x_t;
            const double xc = xt + R * std::sin(th);
            for (int i = 0; i < t.nx; ++i) {
                const double x = t.x[i];
                double z;
                if (x <= xt)
                    z = cfg.prof_z_f + R * (1.0 - std::cos(th)) + (xt - x) * std::tan(th);
                else if (x < xc)
                    z = cfg.prof_z_f + R - std::sqrt(R * R - (x - xc) * (x - xc));
                else
                    z = cfg.prof_z_f;
                for (int j = 0; j < t.ny; ++j) t.z[(size_t)i * t.ny + j] = z;
            }
        }
    }
    t.dx = cfg.dx;
    terrain = std::move(t);
}

double Simulation::releaseHeight(double px, double py) const {
    const RasterData& r = release_raster_;
    if (r.ncols == 0 || r.nrows == 0) return 0.0;
    // node k of the release raster sits at rr_x0_ + k*cellsize in terrain
    // coordinates; the upper bound compares against that absolute last node
    // coordinate (src/complex_terrain.cpp:144-152)
    const double x_back = rr_x0_ + (r.ncols - 1) * r.cellsize;
    const double y_back = rr_y0_ + (r.nrows - 1) * r.cellsize;
    const double xr = px - rr_x0_, yr = py - rr_y0_;
    if (xr < 0 || yr < 0 || xr >= x_back || yr >= y_back) return 0.0;
    const double sx = (rr_x0_ + 1 * r.cellsize) - rr_x0_, sy = (rr_y0_ + 1 * r.cellsize) - rr_y0_;
    const int i = (int)std::floor(xr / sx), j = (int)std::floor(yr / sy);
    if (i < 0 || j < 0 || i >= r.ncols - 1 || j >= r.nrows - 1) return 0.0;
    const double fx = (xr - i * sx) / sx, fy = (yr - j * sy) / sy;
    const size_t n = (size_t)r.nrows;
    const double z00 = r.z[(size_t)i * n + j], z10 = r.z[(size_t)(i + 1) * n + j];
    const double z01 = r.z[(size_t)i * n + j + 1], z11 = r.z[(size_t)(i + 1) * n + j + 1];
    const double nd = r.nodata;
    if (std::fabs(z00 - nd) < 1e-6 || std::fabs(z10 - nd) < 1e-6 || std::fabs(z01 - nd) < 1e-6 ||
        std::fabs(z11 - nd) < 1e-6)
        return 0.0;
    const double z0 = z00 * (1.0 - fx) + z10 * fx;
    const double z1 = z01 * (1.0 - fx) + z11 * fx;
    const double hgt = z0 * (1.0 - fy) + z1 * fy;
    return hgt > 0.0 ? hgt : 0.0;
}

// cos and sin of one angle from one sincos() call, as a GCC build of the
// reference computes the pair (glibc's sincos and cos/sin may differ in the
// last bit)
static void rotate(double px, double py, double ang, double& ox, double& oy) {
    double s, c;
    sincos(ang, &s, &c);
    ox = c * px - s * py;
    oy = s * px + c * py;
}

void Simulation::initializeParticles() {
    if (terrain.nx == 0) throw std::runtime_error("initializeTerrain must run before initializeParticles");
    const double dx = cfg.dx, inv_dx = 1. / dx;
    double x_min = 0, x_max = 0, y_min = 0, y_max = 0, area = 0, volume = 0;
    switch (cfg.release) {
        case ReleaseShapeType::Rectangle: {
            const double diag = std::sqrt(cfg.rect_Lx * cfg.rect_Lx + cfg.rect_Ly * cfg.rect_Ly) / 2.;
            x_min = cfg.rect_cx - diag;
            x_max = cfg.rect_cx + diag;
            y_min = cfg.rect_cy - diag;
            y_max = cfg.rect_cy + diag;
            area = cfg.rect_Lx * cfg.rect_Ly;
            break;
        }
        case ReleaseShapeType::Ellipse: {
            const double big = std::max(cfg.ell_rx, cfg.ell_ry);
            x_min = cfg.ell_cx - big;
            x_max = cfg.ell_cx + big;
            y_min = cfg.ell_cy - big;
            y_max = cfg.ell_cy + big;
            area = M_PI * cfg.ell_rx * cfg.ell_ry;
            break;
        }
        case ReleaseShapeType::Polygon: {
            polygon_ = read_shapefile_polygon(cfg.polygon_path);
            x_min = y_min = std::numeric_limits<double>::max();
            x_max = y_max = std::numeric_limits<double>::lowest();
            for (auto& p : polygon_) {
                p.first -= terrain.xll;
                p.second -= terrain.yll;
                x_min = std::min(x_min, p.first);
                x_max = std::max(x_max, p.first);
                y_min = std::min(y_min, p.second);
                y_max = std::max(y_max, p.second);
            }
            area = polygon_area(polygon_);
            break;
        }
        case ReleaseShapeType::Raster: {
            release_raster_ = read_asc(cfg.raster_path);
            const RasterData& r = release_raster_;
            rr_x0_ = r.xllcorner - terrain.xll;
            rr_y0_ = r.yllcorner - terrain.yll;
            x_min = rr_x0_;
            y_min = rr_y0_;
            x_max = x_min + r.cellsize * (r.ncols - 1);
            y_max = y_min + r.cellsize * (r.nrows - 1);
            const double cell_area = r.cellsize * r.cellsize;
            int64_t count = 0;
            for (int j = 0; j < r.nrows; ++j)
                for (int i = 0; i < r.ncols; ++i) {
                    const double hc = r.z[(size_t)i * r.nrows + j];
                    if (hc >= cfg.raster_h_min && hc > 0.) {
                        ++count;
                        volume += hc * cell_area;
                    }
                }
            area = count * cell_area;
            break;
        }
    }
    if (cfg.release != ReleaseShapeType::Raster) volume = area * cfg.h0;

    x_min = std::floor(x_min * inv_dx) * dx;
    x_max = std::ceil(x_max * inv_dx) * dx;
    y_min = std::floor(y_min * inv_dx) * dx;
    y_max = std::ceil(y_max * inv_dx) * dx;

    std::vector<std::pair<double, double>> samples;
    if (cfg.sampling == SamplingMethod::Uniform) {
        // four particles per cell at the quarter points, i-major then j, then
        // (1/4,1/4), (3/4,1/4), (1/4,3/4), (3/4,3/4)
        static const double ox[4] = {0.25, 0.75, 0.25, 0.75}, oy[4] = {0.25, 0.25, 0.75, 0.75};
        const int ncx = (int)std::round((x_max - x_min) * inv_dx);
        const int ncy = (int)std::round((y_max - y_min) * inv_dx);
        samples.reserve((size_t)std::max(0, ncx) * std::max(0, ncy) * cfg.ppc);
        for (int i = 0; i < ncx; ++i)
            for (int j = 0; j < ncy; ++j)
                for (int k = 0; k < cfg.ppc; ++k)
                    samples.emplace_back(((double)i + ox[k]) * dx, ((double)j + oy[k]) * dx);
    } else {
        samples = poisson_disk(dx, x_max - x_min, y_max - y_min, cfg.ppc);
    }

    ParticleSet ps;
    std::vector<double> hs;
    for (const auto& s : samples) {
        const double px = s.first + x_min, py = s.second + y_min;
        bool keep = false;
        double hk = cfg.h0;
        switch (cfg.release) {
            case ReleaseShapeType::Rectangle: {
                double lx, ly;
                rotate(px - cfg.rect_cx, py - cfg.rect_cy, -cfg.rect_rot, lx, ly);
                keep = std::fabs(lx) <= cfg.rect_Lx / 2. && std::fabs(ly) <= cfg.rect_Ly / 2.;
                break;
            }
            case ReleaseShapeType::Ellipse: {
                double lx, ly;
                rotate(px - cfg.ell_cx, py - cfg.ell_cy, -cfg.ell_rot, lx, ly);
                keep = std::pow(lx / cfg.ell_rx, 2.) + std::pow(ly / cfg.ell_ry, 2.) <= 1.;
                break;
            }
            case ReleaseShapeType::Polygon:
                keep = point_in_polygon(px, py, polygon_);
                break;
            case ReleaseShapeType::Raster:
                hk = releaseHeight(px, py);
                keep = hk >= cfg.raster_h_min && hk > 0.;
                break;
        }
        if (keep) {
            ps.x.push_back(px);
            ps.y.push_back(py);
            hs.push_back(hk);
        }
    }
    const size_t n = ps.x.size();
    double sum_h = 0.0;
    for (size_t p = 0; p < n; ++p) sum_h += hs[p];
    // every particle covers the same planform area, scaled so that
    // sum_p h_p A_p is the release volume (src/simulation.cpp:436-458)
    const double A0 = volume / sum_h;
    ps.h = hs;
    ps.A.assign(n, A0);
    ps.m.resize(n);
    for (size_t p = 0; p < n; ++p) ps.m[p] = cfg.rho * ps.h[p] * ps.A[p];
    ps.release_area = area;
    ps.release_volume = volume;
    double total = 0.0;
    for (size_t p = 0; p < n; ++p) total += ps.h[p] * ps.A[p];
    particles = std::move(ps);
    if (verbose) {
        std::cout << "Total volume: " << total << std::endl;
        std::cout << "Number of particles: " << n << std::endl;
    }
}

// --- driver and writers -----------------------------------------------------

namespace {

struct RunGuard {
    slab_run* r = nullptr;
    ~RunGuard() {
        if (r) slab_run_destroy(r);
    }
};

void check(slab_status st, const char* what) {
    if (st != SLAB_OK) throw std::runtime_error(std::string(what) + ": " + slab_last_error());
}

void save_particles(slab_run* run, int64_t np, const std::string& path) {
    std::vector<double> x(np), y(np), vx(np), vy(np), vz(np), z(np), h(np), A(np);
    check(slab_run_download_particles(run, x.data(), y.data(), vx.data(), vy.data(), vz.data(), z.data(), h.data(),
                                      A.data(), nullptr),
          "download particles");
    write_ply(path, np, x.data(), y.data(), vx.data(), vy.data(), vz.data(), z.data(), h.data(), A.data());
}

// Six bands h, v_norm, m, v_x, v_y, v_z over the remesh window, nodata where
// the float height is zero (src/save_function.cpp:96-226).
void save_grid(slab_run* run, const Terrain& t, const std::string& path) {
    int32_t i0 = 0, j0 = 0, wx = 0, wy = 0;
    check(slab_run_download_grid(run, &i0, &j0, &wx, &wy, nullptr, nullptr, nullptr, nullptr, nullptr), "grid window");
    const size_t n = (size_t)wx * wy;
    std::vector<double> h(n), vx(n), vy(n), vz(n), m(n);
    check(slab_run_download_grid(run, &i0, &j0, &wx, &wy, h.data(), vx.data(), vy.data(), vz.data(), m.data()),
          "download grid");
    const float nodata = -9999.0f;
    std::vector<float> b[6];
    for (auto& v : b) v.resize(n);
    for (size_t k = 0; k < n; ++k) {
        const float hv = (float)h[k];
        if (hv == 0.0f) {
            for (auto& v : b) v[k] = nodata;
        } else {
            b[0][k] = hv;
            b[1][k] = (float)std::sqrt(vx[k] * vx[k] + vy[k] * vy[k]);
            b[2][k] = (float)m[k];
            b[3][k] = (float)vx[k];
            b[4][k] = (float)vy[k];
            b[5][k] = (float)vz[k];
        }
    }
    const double x0 = t.xll + (double)i0 * t.dx;
    const double ytop = t.yll + (double)(j0 + wy - 1) * t.dx;
    write_geotiff(path, wx, wy, x0, ytop, t.dx, {"h", "v_norm", "m", "v_x", "v_y", "v_z"},
                  {b[0].data(), b[1].data(), b[2].data(), b[3].data(), b[4].data(), b[5].data()}, nodata);
}

// Two bands h_max, v_norm_max over the whole terrain (src/save_function.cpp:228-291).
void save_max(slab_run* run, const Terrain& t, const std::string& path) {
    const size_t n = (size_t)t.nx * t.ny;
    std::vector<double> hm(n), vm2(n);
    check(slab_run_download_max(run, hm.data(), vm2.data()), "download max maps");
    const float nodata = -9999.0f;
    std::vector<float> bh(n), bv(n);
    for (size_t k = 0; k < n; ++k) {
        const float hv = (float)hm[k];
        if (hv == 0.0f) {
            bh[k] = bv[k] = nodata;
        } else {
            bh[k] = hv;
            bv[k] = (float)std::sqrt(vm2[k]);
        }
    }
    write_geotiff(path, t.nx, t.ny, t.xll, t.yll + t.y.back(), t.dx, {"h_max", "v_norm_max"}, {bh.data(), bv.data()},
                  nodata);
}

} // namespace

int64_t Simulation::simulate() {
    if (particles.x.empty()) throw std::runtime_error("no particles to simulate");
    if (sim_dir.empty()) throw std::runtime_error("initialize must run before simulate");
    slab_terrain st{terrain.nx, terrain.ny, terrain.dx, terrain.z.data()};
    slab_params sp{cfg.gravity, cfg.rho, cfg.C_cfl, cfg.dt_max, cfg.flip_ratio, cfg.output_dt, cfg.total_time,
                   cfg.precision};
    slab_friction sf{};
    sf.type = (int32_t)cfg.friction;
    sf.mu = cfg.mu;
    sf.xi = cfg.xi;
    sf.mu1 = cfg.mu1;
    sf.mu2 = cfg.mu2;
    sf.I0 = cfg.I0;
    sf.d = cfg.d;
    sf.phi = cfg.phi;
    RunGuard g;
    check(slab_run_create(&st, &sp, &sf, &g.r), "create run");
    const int64_t np = (int64_t)particles.x.size();
    check(slab_run_upload_particles(g.r, np, particles.x.data(), particles.y.data(), nullptr, nullptr, nullptr,
                                    particles.h.data(), particles.A.data(), particles.m.data()),
          "upload particles");
    const int nk = slab_kernel_count();
    if (timing) check(slab_run_timing(g.r, 1, nullptr, nullptr, 0), "timing");

    if (verbose) std::cout << "Starting simulation..." << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    check(slab_run_prime(g.r), "prime");
    int frame = 0;
    save_particles(g.r, np, (fs::path(particle_dir) / frameName("particles_frame_", frame, ".ply")).string());
    save_grid(g.r, terrain, (fs::path(raster_dir) / frameName("rasters_frame_", frame, ".tif")).string());
    const int width = 8;
    if (verbose)
        std::cout << "Time step: " << std::setw(width) << 0 << "    Time: " << 0.0 << " / " << cfg.total_time << " s"
                  << std::endl;
    slab_step_info info{};
    for (;;) {
        check(slab_run_step(g.r, 1 << 30, &info), "step");
        if (info.frame_hit) {
            frame = info.frame;
            save_particles(g.r, np, (fs::path(particle_dir) / frameName("particles_frame_", frame, ".ply")).string());
            save_grid(g.r, terrain, (fs::path(raster_dir) / frameName("rasters_frame_", frame, ".tif")).string());
            if (verbose)
                std::cout << "Time step: " << std::setw(width) << info.step << "    Time: " << std::left
                          << std::setw(10) << info.time << std::right << " / " << cfg.total_time << " s" << std::endl;
        }
        if (info.finished || info.steps_done == 0) break;
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (verbose) {
        std::cout << "Simulation took " << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()
                  << " ms." << std::endl;
        if (timing) {
            std::vector<double> ms(nk);
            std::vector<int64_t> cnt(nk);
            check(slab_run_timing(g.r, 1, ms.data(), cnt.data(), nk), "timing");
            // kernels of one step, grouped by the reference's phases
            auto sum = [&](std::initializer_list<const char*> names) {
                double s = 0;
                for (int k = 0; k < nk; ++k)
                    for (const char* nm : names)
                        if (std::string(slab_kernel_name(k)) == nm) s += ms[k];
                return s / 1e3;
            };
            std::cout << "updateDt took " << sum({"ctl_begin"}) << " s." << std::endl;
            std::cout << "sort took " << sum({"tile_sort"}) << " s." << std::endl;
            std::cout << "remesh took " << sum({"tile_reduce", "tile_scan", "tile_offsets", "scatter"}) << " s."
                      << std::endl;
            std::cout << "P2G took " << sum({"p2g"}) << " s." << std::endl;
            std::cout << "computeGridVelocity took " << sum({"grid"}) << " s." << std::endl;
            std::cout << "G2P took " << sum({"g2p", "ctl_end"}) << " s." << std::endl;
        }
    }
    save_max(g.r, terrain, (fs::path(sim_dir) / "max_values.tif").string());
    return info.step;
}

} // namespace slabhost
