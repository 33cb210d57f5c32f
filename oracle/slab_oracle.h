/*
 * slab_oracle.h -- CPU FP64 restatement of Slab-Pro's depth-averaged MPM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU path is compared
 * against (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).
 * Nothing in the product library links, loads or calls it.
 *
 * Each function follows the reference file:line cited beside it
 * (ALMO-SLF-ETH/Slab-Pro, src/).  Arithmetic is written operation by
 * operation in the reference's evaluation order and built with
 * -ffp-contract=off, so it is reproducible across hosts; the reference's own
 * Release build (-O3 -march=native, GCC default -ffp-contract=fast) may fuse
 * multiply-adds and therefore differs from this restatement in the last ulp.
 *
 * Parity pinning (what checks each part; see tests/test_oracle_*.py):
 *   orc_N / orc_dNdu            closed forms, partition of unity, first moment
 *   terrain precompute          exact plane gradients, last row/col zero
 *   P2G                         UNIFORM t=0 mass identity sum m_i=(1-2^-10) sum m_p,
 *                               brute-force dense evaluation on tiny inputs
 *   grid update                 static Coulomb, sliding-block s(t), Voellmy and
 *                               mu(I) terminal speeds
 *   P2G pressure term           momentum balance of a frictionless dam break
 *                               (1/2 rho g cos^2 h0^2 t per unit width; flat and
 *                               45-degree cross-slope)
 *   G2P                         partition of unity (uniform field), mass invariance,
 *                               sum h_p A_p = sum m_p / rho, height update
 *                               h/(1+(a+b)dt) under a linear velocity field,
 *                               FLIP/PIC weights from velocities alternating
 *                               about an exactly reproduced mean
 *   each of the last three      fails when its term is planted wrong
 *                               (ORC_MUTANT builds, tests/test_oracle_mutants.py)
 *   updateDt / frame snapping   closed form
 *   remesh window               closed form (C0 t=0 window i 34..66, j 54..96)
 *   sort                        permutation + grid-sum invariance
 *   UNIFORM sampler / shapes    closed-form particle counts and areas
 * Full-run trajectories of the reference itself are "parity unpinned": the
 * reference needs Eigen3/GDAL/nlohmann-json >= 3.10.5, none of which exist
 * here, so it cannot be built and no reference output exists.
 */
#ifndef SLAB_ORACLE_H
#define SLAB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* friction law ids = FrictionType order, tools.hpp:37 */
enum { ORC_COULOMB = 0, ORC_VOELLMY = 1, ORC_MU_I = 2 };

typedef struct {
    double gravity, rho, C_cfl, dt_max, flip_ratio, output_dt, total_time;
    int32_t fric_type;
    double mu, xi, mu1, mu2, I0, d, phi;
} orc_params;

/* Quadratic B-spline and its derivative, tools.hpp:119-140. */
double orc_N(double u);
double orc_dNdu(double u);

/* Terrain: z in ind(i,j,Ny)=i*Ny+j order (tools.hpp:172).  Computes gradients
 * and theta as complex_terrain.cpp:84-111. */
orc_sim* orc_create(int32_t nx, int32_t ny, double dx, const double* z, const orc_params* prm);
void orc_destroy(orc_sim* s);

/* Particles as initializeParticles leaves them (simulation.cpp:443-466):
 * v = vz = 0, mass = rho*h*A, theta from the floor cell; z by the 4x4 wip
 * interpolation at simulation.cpp:458-465.  m may be NULL (then rho*h*A). */
int orc_set_particles(orc_sim* s, int32_t np, const double* x, const double* y,
                      const double* h, const double* A, const double* m);

/* simulate() lines 493-495: remesh, P2G, computeGridVelocity with dt = 0. */
void orc_prime(orc_sim* s);
/* One pass of the while loop at simulation.cpp:500-510: advanceStep, time += dt,
 * frame check.  Returns 1 when a frame was hit.  Returns -1 once time >= total. */
int orc_step(orc_sim* s);
/* Run steps until time >= total_time or max_steps; returns steps run. */
int64_t orc_run(orc_sim* s, int64_t max_steps);

/* Scalars: out[0]=time, out[1]=dt, out[2]=frame, out[3]=time_step,
 * out[4..9]=phase seconds (updateDt, remesh, sort, P2G, grid, G2P). */
void orc_scalars(const orc_sim* s, double out[10]);
/* Window of the current grid (simulation.cpp:608-612). */
void orc_window(const orc_sim* s, int32_t out[4]); /* iL, jB, Nx, Ny */
/* Grid fields over the window in ind(i,j,Ny_window) order.  Any may be NULL. */
void orc_get_grid(const orc_sim* s, double* mass, double* vx, double* vy, double* vz,
                  double* h, double* flipx, double* flipy, double* Fx, double* Fy);
/* Particles in ORIGINAL (creation) order.  Any may be NULL. */
void orc_get_particles(const orc_sim* s, double* x, double* y, double* vx, double* vy,
                       double* vz, double* z, double* h, double* A, double* m);
/* Overwrite particle velocity (original order); used by closed-form tests. */
void orc_set_velocity(orc_sim* s, const double* vx, const double* vy);
/* Full-terrain maps h_max, v_max2 and terrain derived fields (NULL to skip). */
void orc_get_max(const orc_sim* s, double* h_max, double* v_max2);
void orc_get_terrain(const orc_sim* s, double* gx, double* gy, double* theta);
/* Disable the every-50-steps block sort (simulation.cpp:540) for tests that
 * want the unsorted order; default on. */
void orc_set_sort(orc_sim* s, int enabled);
void orc_set_time(orc_sim* s, double time, int32_t frame);

/* ---- front-end restatements (initializeParticles, simulation.cpp:259-466) ---- */
/* UNIFORM 2x2 sampler over the dx-snapped bounding box (simulation.cpp:355-385)
 * filtered by a rotated rectangle (kind 0) or ellipse (kind 1) as at 397-409.
 * Writes up to cap positions; returns the count (may exceed cap: call again). */
int64_t orc_uniform_shape(double dx, int32_t kind, double cx, double cy, double a, double b,
                          double rot_rad, double* x, double* y, int64_t cap);
/* Same sampler filtered by the even-odd polygon test of polygon_tools.cpp:87-101. */
int64_t orc_uniform_polygon(double dx, const double* vx, const double* vy, int32_t nv,
                            double* x, double* y, int64_t cap);
/* Bilinear release-height lookup of complex_terrain.cpp:132-183 on a raster
 * with node spacing cs, origin (ox, oy) in terrain-local coordinates. */
double orc_raster_height(const double* z, int32_t ncols, int32_t nrows, double ox, double oy,
                         double cs, double nodata, double px, double py);

#ifdef __cplusplus
}
#endif
#endif
