// slab_engine32.hpp -- the FP32 step engine: one or many scenarios ("slots")
// on one shared DEM, advanced together by one launch sequence per step.
//
// A single FP32 run is a batch of one slot.  Each slot has its own friction
// law and parameters, simulated time, dt (CFL of its own particles), frame
// counter and total_time (the per-run state of src/simulation.cpp:484-582),
// its own max maps and grid.  Tile ids are slot * ntd + DEM tile, so the
// shared sort / scan machinery orders particles by (slot, tile, cell,
// creation index) and one P2G / grid / G2P launch covers every slot.
#pragma once
#include "slab_abi.h"
#include "slab_common.hpp"

#include <string>
#include <vector>

namespace slab {

enum SlotStatus { SLOT_FREE = 0, SLOT_PRIMING = 1, SLOT_RUNNING = 2, SLOT_DONE = 3, SLOT_FAILED = 4 };

constexpr int VMAX_SHARDS = 256; // per-slot shards of the max |v|^2 atomics

// Per-slot control block (device).
struct Slot {
    double time, dt, total_time, output_dt, dt_max, C_cfl, flip_ratio, rho, gravity;
    double mu, xi, mu1, mu2, I0, d, phi;
    int32_t fric, status, frame, step;
    int32_t out_step, frame_hit, keep_grid, np;
    int32_t last_out_epoch, err, steps_in_call;
    int32_t vz_due; // this step's G2P refreshes particle vz (an output step may follow)
};

// FP32 particle state (structure of arrays).  Position = floor cell packed as
// (ci << 16) | cj plus an in-cell offset (ox, oy) in [0,1) cell units.
struct PB32 {
    int32_t* cell;
    float2* o;            // offsets in the floor cell (cell units), x and y in one 8-byte word
    float2* v;            // vx, vy
    float *vz, *h, *m, *z;
    int32_t *pid, *slot;
};

// Outflow: a particle whose position leaves the cells [1, N-3] of either axis
// (where its 3x3 stencil would leave the terrain) is parked for the rest of
// its run: its state is frozen (velocity zero), it no longer enters P2G / G2P
// or the CFL dt.  Marked by this bit in the particle's slot field.
constexpr int32_t SLOT_OUT = 1 << 30;
constexpr int32_t SLOT_MASK = SLOT_OUT - 1;

constexpr int REGION = TILE + 2;              // 18x18 P2G partial region per particle tile
constexpr int REGION_NODES = REGION * REGION;
constexpr int P2G_FIELDS = 7;                 // m, P_x, P_y, H, F_x, F_y, P_z
constexpr int P2G_PZ = P2G_FIELDS - 1;        // P_z last: only output steps store it

// Global control of the engine (device); extends the per-step fields of Ctl.
struct ECtl {
    int32_t nx, ny, ntx, nty, ntd, nslots, single, has_out;
    int32_t epoch, parity, p2g_parity, stop;
    int32_t err, n_store, n_sorted, n_ptiles;
    int32_t n_gtiles, tile_cap, steps_in_call, n_active;
    int32_t compact;       // this pass only re-sorts (drops released particles)
    int32_t reorder;       // this step's G2P writes the other buffer in sorted order
    int32_t reorder_every; // storage re-sort period in steps (other steps update in place)
    int32_t pend_end;      // a started step still has to be ended (time, frames, buffer flip)
    int32_t gtile_cap;     // grid tiles the grid-velocity pool holds (batches; all tiles for a single run)
    int32_t hm_cap;        // tiles the max-map pool holds (batches)
    int32_t hm_top;        // max-map pool tiles in use (the free ones are hm_stack[hm_top..hm_cap))
    int32_t pad_;
    double dx, inv_dx;
};

struct Engine32;

// Message of every call refused after a step stopped on an error.
constexpr const char* kStoppedByError = "stopped by an earlier error (a tile capacity was exceeded); destroy it";

// Host-side engine API used by the C ABI (slab_run_* with SLAB_FP32, slab_batch_*).
struct EngineResult {
    int ok;
    std::string msg;
};

Engine32* engine_create(int nx, int ny, double dx, const double* z, int nslots, int64_t max_particles, bool single,
                        std::string* err);
void engine_destroy(Engine32* e);
// Adds a scenario; returns slot id >= 0, or -1 (no free slot / no capacity) with msg.
int engine_admit(Engine32* e, const Slot& params, int64_t np, const double* x, const double* y, const double* vx,
                 const double* vy, const double* vz, const double* h, const double* A, const double* m,
                 std::string* err);
// Runs up to max_steps global steps (stops early in single mode at frames or at the end).
bool engine_step(Engine32* e, int max_steps, std::string* err);
bool engine_pull(Engine32* e, std::string* err);              // refresh host mirrors of ECtl and slots
const ECtl& engine_ctl(const Engine32* e);
const Slot& engine_slot(const Engine32* e, int s);
bool engine_release(Engine32* e, int slot, std::string* err); // frees a slot (its particles drop out)
bool engine_particles(Engine32* e, int slot, double* const dst[9], std::string* err);
// The same download staged through the library's fixed per-device scratch
// instead of allocations (slots of at most SLAB_NOALLOC_MAX_PARTICLES).
bool engine_particles_noalloc(Engine32* e, int slot, double* const dst[9], std::string* err);
bool engine_max_maps(Engine32* e, int slot, double* h_max, double* v_max2, std::string* err);
bool engine_window(Engine32* e, int slot, int* i0, int* j0, int* wnx, int* wny, double* const dst[5],
                   std::string* err);
bool engine_diag(Engine32* e, int slot, double out[8], std::string* err);
bool engine_set_velocity(Engine32* e, int slot, const double* vx, const double* vy, std::string* err);
bool engine_set_time(Engine32* e, int slot, double time, int frame, std::string* err);
bool engine_keep_grid(Engine32* e, int slot, int on, std::string* err);
void engine_timing(Engine32* e, int enable, double* ms, int64_t* counts, int n);
int64_t engine_particle_bytes();                      // device bytes per particle of capacity
int64_t engine_slot_bytes(const Engine32* e);         // device bytes per slot (max maps; dense grid of a single run)
// Active-tile capacity and use: out = {grid-tile capacity, grid tiles of the
// last step, particle-tile capacity, particle tiles of the last step,
// device bytes per grid tile, device bytes per particle tile, max-map tile
// capacity, max-map tiles in use, device bytes per max-map tile}.
void engine_tiles(const Engine32* e, int64_t out[9]);
// Admission reserves this many times a release's tiles for its flow to spread.
bool engine_set_tile_growth(Engine32* e, int factor, std::string* err);
int64_t engine_free_particles(const Engine32* e);     // capacity left for admissions
int engine_use_graph(Engine32* e, int on);             // capture the step sequence in a hipGraph
// Storage re-sort period: every k-th step G2P writes the particles in sorted
// order to the other buffer (dropping released ones); other steps update them
// in place.  Results do not depend on k (the computation order is canonical).
bool engine_set_reorder(Engine32* e, int k, std::string* err);
// Times candidate launch shapes (storage re-sort period, workgroups of the
// sort + P2G kernel, in-LDS capacity of the tile sort) on the next steps of the
// live run and keeps the fastest; out = {period, workgroups, sort capacity}.
// The steps it runs are ordinary steps.
bool engine_autotune(Engine32* e, int trial_steps, int out[3], std::string* err);
// In-LDS capacity of the tile sort: 1024, 2048 or 4096 particles (default
// 4096).  Results do not depend on it.
bool engine_set_sort_capacity(Engine32* e, int cap, std::string* err);
// The current launch shape {re-sort period, workgroups, sort capacity}.
void engine_launch_shape(const Engine32* e, int out[3]);

} // namespace slab
