/* slab_frontend.h -- C ABI of the configuration front end in libslabpro.so.
 *
 * The reference's executable is driven by one JSON file
 * (src/main.cpp:29-58: loadFromJSON, initialize, initializeTerrain,
 * initializeParticles, simulate).  These entry points expose the same
 * sequence, plus the file readers/writers and the samplers it uses, so a
 * host program (or a test without a GPU) can call each stage.  Every
 * function except slab_sim_simulate runs on the CPU only.  Errors return a
 * slab_status and leave a message for slab_last_error().
 */
#ifndef SLAB_FRONTEND_H
#define SLAB_FRONTEND_H

#include <stdint.h>

#include "slab_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slab_sim slab_sim;

/* Simulation::loadFromJSON (src/simulation.cpp:118-208).  Relative paths in
 * the file resolve against the working directory, as in the reference.
 * Optional key numerical_params.precision: "FP64" (default) or "FP32". */
slab_status slab_sim_load_json(const char* json_path, slab_sim** out);

/* Simulation::initialize (src/simulation.cpp:54-81 with
 * createSimulationDirectories, src/tools.cpp:28-51): creates
 * <project_dir>/<sim_name>/{rasters,particles}, sets nFiles/padding and
 * copies the JSON file into the simulation folder. */
slab_status slab_sim_initialize(slab_sim* s);

/* Simulation::initializeTerrain (src/simulation.cpp:210-257): ESRI ASCII DEM
 * or the analytical 5-degree plane. */
slab_status slab_sim_initialize_terrain(slab_sim* s);

/* Simulation::initializeParticles (src/simulation.cpp:259-476): release
 * shape, UNIFORM or POISSON sampling, heights, areas, masses. */
slab_status slab_sim_initialize_particles(slab_sim* s);

/* Simulation::simulate (src/simulation.cpp:484-522) on the GPU: writes
 * particles/particles_frame_<k>.ply and rasters/rasters_frame_<k>.tif for
 * the t = 0 frame and every output frame, then max_values.tif.  steps may
 * be NULL. */
slab_status slab_sim_simulate(slab_sim* s, int64_t* steps);

/* Console output of the stages (default on) and per-phase kernel timing
 * printed after the run (default off). */
slab_status slab_sim_set_verbose(slab_sim* s, int32_t verbose, int32_t timing);
/* Overrides the precision read from the JSON file (slab_precision). */
slab_status slab_sim_set_precision(slab_sim* s, int32_t precision);

/* Terrain after slab_sim_initialize_terrain; z (nx*ny, z[i*ny+j]) may be NULL. */
slab_status slab_sim_terrain(const slab_sim* s, int32_t* nx, int32_t* ny, double* dx, double* xllcorner,
                             double* yllcorner, double* z);
/* Particles after slab_sim_initialize_particles; any array may be NULL. */
int64_t slab_sim_num_particles(const slab_sim* s);
slab_status slab_sim_particles(const slab_sim* s, double* x, double* y, double* h, double* A, double* m);
/* Release area and volume computed by slab_sim_initialize_particles. */
slab_status slab_sim_release(const slab_sim* s, double* area, double* volume);
/* Simulation folder, number of output frames and the frame-number padding
 * set by slab_sim_initialize (src/simulation.cpp:59-60). */
slab_status slab_sim_output(const slab_sim* s, char* sim_dir, int32_t cap, int32_t* n_files, int32_t* padding);
void slab_sim_destroy(slab_sim* s);

/* Poisson-disk samples of [0,Lx]x[0,Ly] (src/poisson_disk_sampling.hpp),
 * from the rand() stream of a fresh process.  Returns the sample count;
 * writes at most cap samples (x, y may be NULL to query the count), or -1. */
int64_t slab_poisson_disk(double dx, double Lx, double Ly, int32_t ppc, double* x, double* y, int64_t cap);

/* Polygon area (OGRPolygon::get_Area as src/polygon_tools.cpp:67-85 calls it)
 * and the point-in-polygon test of src/polygon_tools.cpp:87-100. */
double slab_polygon_area(const double* x, const double* y, int32_t n);
int32_t slab_point_in_polygon(double px, double py, const double* x, const double* y, int32_t n);

/* First polygon's outer ring of an ESRI shapefile (src/polygon_tools.cpp:22-65).
 * *n receives the vertex count; at most cap vertices are written. */
slab_status slab_read_shapefile_polygon(const char* path, double* x, double* y, int32_t cap, int32_t* n);

/* Binary little-endian PLY with float properties x y vx vy vz z h A
 * (src/save_function.cpp:36-76). */
slab_status slab_write_ply(const char* path, int64_t np, const double* x, const double* y, const double* vx,
                           const double* vy, const double* vz, const double* z, const double* h, const double* A);

/* Float32 GeoTIFF, bands[b][i*ny+j] in grid node order, written north-up with
 * geotransform (x0, dx, 0, ytop, 0, -dx), band names and nodata
 * (src/save_function.cpp:96-291). */
slab_status slab_write_geotiff(const char* path, int32_t nx, int32_t ny, double x0, double ytop, double dx,
                               int32_t nbands, const char* const* names, const float* const* bands, float nodata);

#ifdef __cplusplus
}
#endif

#endif
