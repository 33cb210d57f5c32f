/*
 * slab_abi.h -- C ABI of the MI355X depth-averaged MPM step (libslabpro.so).
 *
 * The reference (ALMO-SLF-ETH/Slab-Pro) has no FFI: its step is the set of
 * private member functions of Simulation<T,TV2> (src/simulation.hpp:281-286)
 * driven by Simulation::simulate() (src/simulation.cpp:484-522).  These entry
 * points replace that loop body one for one; each declaration cites the
 * reference code it stands in for.
 *
 * Conventions
 *  - Every call returns slab_status (0 = OK).  No C++ exception crosses the
 *    ABI; slab_last_error() returns a thread-local message for the last
 *    failing call on this thread.
 *  - Host buffers belong to the caller and are copied in/out.  Arrays use the
 *    reference's layouts: node arrays in ind(i,j,Ny) = i*Ny + j order
 *    (src/tools.hpp:172-174), particle arrays in creation order.
 *  - A handle owns its device memory and one HIP stream; it is used by one
 *    host thread at a time.  Different handles may be used concurrently.
 *  - Nothing falls back to the CPU: without a usable gfx950 device,
 *    slab_run_create fails with SLAB_EDEVICE.
 */
#ifndef SLAB_ABI_H
#define SLAB_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slab_run slab_run;

typedef enum {
    SLAB_OK = 0,
    SLAB_EINVAL = 1,  /* bad argument */
    SLAB_ENOMEM = 2,  /* device allocation failed */
    SLAB_EDEVICE = 3, /* no device / HIP runtime error */
    SLAB_EDOMAIN = 4, /* input outside what the step supports */
    SLAB_ENAN = 5,    /* non-finite velocity detected on the device */
    SLAB_ESTATE = 6   /* call not valid in the handle's current state */
} slab_status;

typedef enum { SLAB_FP64 = 0, SLAB_FP32 = 1 } slab_precision;

/* = enum class FrictionType order, src/tools.hpp:37 */
typedef enum { SLAB_COULOMB = 0, SLAB_VOELLMY = 1, SLAB_MU_I = 2 } slab_friction_type;

/* = struct Friction and its members, src/tools.hpp:62-84 */
typedef struct {
    int32_t type;          /* slab_friction_type */
    double mu;             /* Coulomb mu or Voellmy mu */
    double xi;             /* Voellmy xi [m s^-2] */
    double mu1, mu2, I0, d, phi; /* mu(I) */
} slab_friction;

/* = the scalar members read by the step (src/simulation.hpp:288-310) */
typedef struct {
    double gravity, rho, C_cfl, dt_max, flip_ratio, output_dt, total_time;
    int32_t precision; /* slab_precision */
} slab_params;

/* = Grid "terrain" as initializeTerrain leaves it (src/simulation.cpp:210-257):
 * nodes X_i = i*dx, Y_j = j*dx; z[i*ny + j]. */
typedef struct {
    int32_t nx, ny;
    double dx;
    const double* z;
} slab_terrain;

/* Progress of the driver (src/simulation.cpp:500-510 state variables). */
typedef struct {
    double time;       /* simulated time after the last step */
    double dt;         /* dt of the last step */
    int32_t frame;     /* output frame counter */
    int32_t step;      /* time_step counter */
    int32_t frame_hit; /* last step reached output_dt*(frame) (frame already advanced) */
    int32_t finished;  /* time >= total_time */
    int32_t steps_done;/* steps executed by this call */
    int32_t reserved;
} slab_step_info;

/* Creates a run on the current HIP device: uploads the terrain and computes
 * its gradient / slope tables.  Replaces initializeTerrain's tail,
 * terrainGradients + terrainTheta (src/complex_terrain.cpp:84-111), and the
 * zeroing of h_max / v_max2 (src/simulation.cpp:254-256). */
slab_status slab_run_create(const slab_terrain* terrain, const slab_params* params,
                            const slab_friction* friction, slab_run** out);

/* Uploads particles as initializeParticles leaves them
 * (src/simulation.cpp:443-466).  vx, vy, vz may be NULL (zero).  m may be
 * NULL (then rho*h*A as at :452).  Particle order is the creation order.
 * Non-finite inputs are rejected (SLAB_EINVAL). */
slab_status slab_run_upload_particles(slab_run* run, int64_t np, const double* x, const double* y,
                                      const double* vx, const double* vy, const double* vz,
                                      const double* h, const double* A, const double* m);

/* t = 0 priming: remesh, P2G, computeGridVelocity with dt = 0
 * (src/simulation.cpp:493-495).  Seeds h_max with the initial height. */
slab_status slab_run_prime(slab_run* run);

/* Runs up to max_steps passes of the while loop at src/simulation.cpp:500-510
 * (advanceStep = updateDt, remesh, sort, P2G, computeGridVelocity, G2P;
 * time += dt; frame check).  Returns early after a step that hits an output
 * frame or when time >= total_time, so the caller can write outputs.
 * info may be NULL. */
slab_status slab_run_step(slab_run* run, int32_t max_steps, slab_step_info* info);

/* Particles in creation order, as saveParticleData reads them
 * (src/simulation.cpp via src/save_function.cpp:36-76).  Any pointer may be NULL.
 * z, the bed elevation G2P interpolates at each particle (simulation.cpp:930),
 * and vz = v . grad z (:933) are read by the reference only when a frame is
 * saved (and vz by the P2G of such a step): the FP32 form refreshes z on
 * output steps only (frames, the prime, every step after
 * slab_run_keep_grid(1)) and vz on those and on a step after which the next
 * can reach its frame, and carries them unchanged otherwise, so between
 * frames they hold the last such step's values. */
slab_status slab_run_download_particles(slab_run* run, double* x, double* y, double* vx, double* vy,
                                        double* vz, double* z, double* h, double* A, double* m);

/* Grid of the last step over the reference's remesh window
 * (src/simulation.cpp:594-612): i0 = iL, j0 = jB, nx/ny = window size.
 * Call with all field pointers NULL to query the window first.  Field arrays
 * (nx*ny, ind order) as saveGridData reads them (src/save_function.cpp:96-226).
 * Returns SLAB_ESTATE if the last step did not keep grid outputs (only
 * output-frame steps, the prime and runs with slab_run_keep_grid(1) do). */
slab_status slab_run_download_grid(slab_run* run, int32_t* i0, int32_t* j0, int32_t* nx, int32_t* ny,
                                   double* h, double* vx, double* vy, double* vz, double* m);

/* Keep grid output fields after every step (default: only output frames). */
slab_status slab_run_keep_grid(slab_run* run, int32_t enable);

/* Full-terrain h_max and v_max2 maps (src/save_function.cpp:228-291 inputs). */
slab_status slab_run_download_max(slab_run* run, double* h_max, double* v_max2);

/* FP64 diagnostics of the last step: out[0]=sum m_i (grid), out[1]=sum m_i vx_i,
 * out[2]=sum m_i vy_i, out[3]=sum m_p, out[4]=sum h_p A_p, out[5]=dt, out[6]=max|v_p|,
 * out[7]=number of grid nodes with m > 1e-6.  (No reference counterpart: the
 * reference only prints the initial volume, src/simulation.cpp:469-475.) */
slab_status slab_run_diag(slab_run* run, double out[8]);

/* Event timing of each step kernel on the run's stream.  enable=1 starts
 * accumulating; ms[k] / counts[k] receive totals per kernel (k < n, names by
 * slab_kernel_name).  No reference counterpart: it replaces the per-phase
 * timers of advanceStep (src/simulation.cpp:526-562, printed at 512-519). */
slab_status slab_run_timing(slab_run* run, int32_t enable, double* ms, int64_t* counts, int32_t n);
const char* slab_kernel_name(int32_t k);
int32_t slab_kernel_count(void);

/* Overrides the particle velocity (creation order) -- used by closed-form tests
 * and drivers that restart from saved state (setParticles, simulation.hpp:186). */
slab_status slab_run_set_velocity(slab_run* run, const double* vx, const double* vy);

/* Overrides time and frame (setFrame/setDt, simulation.hpp:100-113). */
slab_status slab_run_set_time(slab_run* run, double time, int32_t frame);

const char* slab_last_error(void);
void slab_run_destroy(slab_run* run);

/* Library / device facts: out[0]=device id, out[1]=CU count, out[2]=is gfx950. */
slab_status slab_device_info(int32_t out[3]);

/* Streaming kernels of known size on the current device, the practical HBM
 * ceiling for the step's access patterns.  kind 0: copy n bytes (float4 reads
 * and writes); 1: read n bytes; 2: write n bytes; 3: the particle streams of
 * G2P for n particles (28 B read, 32 B written in place, one per thread);
 * 4 and 5: the copy and the write with non-temporal accesses and 128 B in
 * flight per thread; 6: eight dependent copy phases of n bytes (a -> b,
 * b -> a, ...), one launch each; 7: the same eight phases in one cooperative
 * launch, separated by grid-wide barriers with device-scope release/acquire
 * (6 and 7 use the same workgroup count, four per CU: their difference is the
 * cost of a kernel boundary against a grid barrier); 8 and 9: 6 and 7 with
 * one workgroup per CU.
 * *ms = mean time of iters launches (of the eight launches for kind 6; HIP
 * events).  No reference counterpart: measurement only. */
slab_status slab_stream_probe(int32_t kind, int64_t n, int32_t iters, double* ms);

/* ---------------------------------------------------------------------------
 * Batches: many scenarios on one shared DEM, advanced together (FP32 form).
 * No reference counterpart -- the reference runs one scenario per process
 * (src/main.cpp:29-58).  Each slot is one Simulation's step state: its own
 * friction (src/tools.hpp:62-84), time / dt / frame / total_time
 * (src/simulation.cpp:484-582) and max maps.  Slots advance by their own dt;
 * a slot that reaches total_time is parked (its particles are kept but not
 * computed) until released.  The batch owns one HIP stream.
 * ------------------------------------------------------------------------- */
typedef struct slab_batch slab_batch;

enum { SLAB_SLOT_FREE = 0, SLAB_SLOT_PRIMING = 1, SLAB_SLOT_RUNNING = 2, SLAB_SLOT_DONE = 3, SLAB_SLOT_FAILED = 4 };

typedef struct {
    double time, dt;
    int32_t status, frame, step, np;
    int32_t frame_hit, err;
} slab_slot_info;

/* Shared terrain (complex_terrain.cpp:84-111 precompute) + capacity for
 * max_slots scenarios and max_particles particles in total. */
slab_status slab_batch_create(const slab_terrain* terrain, int32_t max_slots, int64_t max_particles,
                              slab_batch** out);
/* Admits one scenario (particles as initializeParticles leaves them; vx, vy
 * and m may be NULL); it is primed (dt = 0 pass, simulation.cpp:493-495) by
 * the next step.  SLAB_ENOMEM when no slot, particle capacity or active-tile
 * capacity (see slab_batch_tiles) is free, SLAB_EINVAL for non-finite inputs
 * or parameters; *slot = -1 then. */
slab_status slab_batch_admit(slab_batch* b, const slab_params* params, const slab_friction* friction, int64_t np,
                             const double* x, const double* y, const double* vx, const double* vy, const double* h,
                             const double* A, const double* m, int32_t* slot);
/* Runs up to max_steps batch steps (every active slot advances one of its own
 * steps per batch step); returns early when no slot is active.
 * SLAB_EDOMAIN when the particle tiles of the active slots exceed the batch's
 * active-tile capacity: that step is left half started, so from then on
 * slab_batch_step, slab_batch_admit and slab_batch_release return SLAB_ESTATE
 * (downloads still read the last completed step); destroy the batch.  The
 * same holds for slab_run_step of an FP32 run. */
slab_status slab_batch_step(slab_batch* b, int32_t max_steps, int32_t* n_active);
slab_status slab_batch_slot_info(slab_batch* b, int32_t slot, slab_slot_info* out);
/* status[max_slots] and time[max_slots] of every slot (either may be NULL). */
slab_status slab_batch_statuses(slab_batch* b, int32_t* status, double* time);
slab_status slab_batch_download_max(slab_batch* b, int32_t slot, double* h_max, double* v_max2);
/* As slab_run_download_particles for one slot (z as refreshed on the slot's
 * output steps; vz every step). */
slab_status slab_batch_download_particles(slab_batch* b, int32_t slot, double* x, double* y, double* vx, double* vy,
                                          double* vz, double* z, double* h, double* A, double* m);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, code between graph replays, harnesses that forbid
 * hipMalloc): the same outputs as slab_batch_download_particles, staged
 * through a fixed scratch array of the library (a counter plus 9 FP64 fields
 * of SLAB_NOALLOC_MAX_PARTICLES particles, about 18.9 MB, held by every GPU the
 * library runs on until the process ends) instead of per-call allocations;
 * no hipMalloc / hipFree is made.  SLAB_EDOMAIN for a slot of more than
 * SLAB_NOALLOC_MAX_PARTICLES particles.  All calls on one GPU share that GPU's
 * scratch and each overwrites it: calls that overlap on one GPU (different
 * host threads or batches) race. */
#define SLAB_NOALLOC_MAX_PARTICLES 262144
slab_status slab_batch_download_particles_noalloc(slab_batch* b, int32_t slot, double* x, double* y, double* vx,
                                                  double* vy, double* vz, double* z, double* h, double* A,
                                                  double* m);
/* Frees a slot at once, whatever its status (a running scenario is
 * cancelled): its particles are dropped from the device state by one
 * compaction pass inside this call, so their capacity and the slot are free
 * for the next admission.  The other slots are not advanced. */
slab_status slab_batch_release(slab_batch* b, int32_t slot);
/* out[0] = device bytes per particle, out[1] = device bytes per slot (its
 * control block and tile map; its max maps are counted in tiles),
 * out[2] = free particle capacity, out[3] = free slots.
 * With slab_batch_tiles this is the batch's device footprint: particles +
 * slots + active tiles. */
slab_status slab_batch_footprint(slab_batch* b, int64_t out[4]);
/* Active tiles of the batch.  The grid velocity of a step is kept only for its
 * grid tiles (16 x 16 nodes holding particles, and their right / upper /
 * diagonal neighbours), in a pool of fixed capacity; P2G keeps a partial
 * region per particle tile; each slot's h_max / v_max2 maps are kept only for
 * the tiles its flow has touched, in a third pool (tiles return to it when the
 * slot is released).  out[0] = grid-tile capacity, out[1] = grid tiles of the
 * last step, out[2] = particle-tile capacity, out[3] = particle tiles of the
 * last step, out[4] = device bytes per grid tile, out[5] = device bytes per
 * particle tile, out[6] = max-map tile capacity, out[7] = max-map tiles in
 * use, out[8] = device bytes per max-map tile.  No reference counterpart (the
 * reference sizes one bounding-box window per run, src/simulation.cpp:584-642,
 * and one full-terrain pair of max maps). */
slab_status slab_batch_tiles(slab_batch* b, int64_t out[9]);
/* Admission reserves `factor` times the particle and grid tiles of a release
 * (default 4) for its flow to spread; slab_batch_admit refuses a scenario
 * whose reservation does not fit beside the live ones.  factor = 0 admits
 * without a tile check (a step whose tiles then overflow the capacity returns
 * SLAB_EDOMAIN, see slab_batch_step). */
slab_status slab_batch_set_tile_growth(slab_batch* b, int32_t factor);
/* Replay the step's launch sequence from a captured hipGraph (on = 1). */
slab_status slab_batch_use_graph(slab_batch* b, int32_t on);
/* Storage re-sort period of the FP32 form (default 16): every `steps`-th step
 * G2P writes the particles to the other buffer in canonical order (dropping
 * released ones); the steps between update them in place.  Results do not
 * depend on it (every sum runs in canonical order); speed does.  No reference
 * counterpart beyond the every-50-steps block sort of src/simulation.cpp:540. */
slab_status slab_batch_set_resort(slab_batch* b, int32_t steps);
/* The same for a single FP32 run (after slab_run_upload_particles); a no-op
 * for the FP64 form, which re-sorts its storage every step. */
slab_status slab_run_set_resort(slab_run* run, int32_t steps);
/* Autotuner of the FP32 form: times candidate launch shapes -- storage
 * re-sort period {1, 4, 16, 64}, then workgroups of the tile sort + P2G kernel
 * {1024, 2048, 4096, 8192}, then the in-LDS capacity of the tile sort {1024,
 * 2048, 4096} -- on the next steps of the live run (each trial runs
 * max(trial_steps, 2 x period) ordinary steps; results do not depend on the
 * shape) and keeps the fastest.  out[0] = period, out[1] = workgroups; the
 * whole shape is read back with slab_*_launch_shape.  The run variant needs
 * slab_run_prime first and is a no-op for the FP64 form.  No reference
 * counterpart (its sort block is fixed at 4x4 cells, src/simulation.cpp:645). */
slab_status slab_batch_autotune(slab_batch* b, int32_t trial_steps, int32_t out[2]);
slab_status slab_run_autotune(slab_run* run, int32_t trial_steps, int32_t out[2]);
/* In-LDS capacity of the FP32 tile sort, in particles: 1024, 2048 or 4096
 * (default).  Tiles above it are ordered by cell in global memory and then in
 * rounds of whole cells of at most half the capacity.  A smaller capacity holds
 * fewer registers and less LDS per workgroup (more workgroups per CU); results
 * do not depend on it.  A no-op for the FP64 form.  No reference counterpart
 * (src/simulation.cpp:644-690 sorts by 4x4-cell blocks on the host). */
slab_status slab_batch_set_sort_capacity(slab_batch* b, int32_t particles);
slab_status slab_run_set_sort_capacity(slab_run* run, int32_t particles);
/* The current FP32 launch shape: out[0] = storage re-sort period, out[1] =
 * workgroups of the tile sort + P2G kernel, out[2] = sort capacity (the FP64
 * form reports {1, 0, 0}; the grid kernel's workgroups follow the last step's
 * grid-tile count).  No reference counterpart. */
slab_status slab_batch_launch_shape(slab_batch* b, int32_t out[3]);
slab_status slab_run_launch_shape(slab_run* run, int32_t out[3]);
slab_status slab_batch_timing(slab_batch* b, int32_t enable, double* ms, int64_t* counts, int32_t n);
void slab_batch_destroy(slab_batch* b);

#ifdef __cplusplus
}
#endif
#endif
