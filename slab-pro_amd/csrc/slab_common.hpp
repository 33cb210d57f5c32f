// slab_common.hpp -- shared device/host definitions of the MI355X MPM step.
//
// Node layout: every node array of the run covers the whole terrain in the
// reference's ind(i,j,Ny) = i*Ny + j order (src/tools.hpp:172-174), so uploads
// and downloads need no reordering.  Work is organised in 16x16-node tiles:
// tile (tx,ty) = (i>>4, j>>4), linear id tx*nty + ty.  A particle's "cell" is
// the base node (ib,jb) of its 3x3 quadratic stencil (src/simulation.cpp:741-742);
// its sort key is tile*256 + ((ib&15)<<4 | (jb&15)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slab {

constexpr int TILE = 16;
constexpr int TILE_CELLS = TILE * TILE; // 256
constexpr int CELL_STRIDE = TILE_CELLS + 1; // per-tile cell_begin entries (last = tile end)
constexpr int WG = 256;
constexpr int SCAN_ITEMS = 4; // tiles per thread in the tile scan
constexpr int SCAN_BLOCK = WG * SCAN_ITEMS;
constexpr int PART_SHARDS = 64; // shards of the FP64 {bbox, max |v|^2} partials (5 x u64 each)

enum StopCode { RUNNING = 0, STOP_FRAME = 1, STOP_DONE = 2, STOP_ERROR = 3 };

// Device control block: parameters plus the driver state of
// src/simulation.cpp:484-582, kept on the device so steps need no host sync.
struct Ctl {
    // parameters (src/simulation.hpp:288-310, src/tools.hpp:62-84)
    double gravity, rho, C_cfl, dt_max, flip_ratio, output_dt, total_time, dx, inv_dx;
    double mu, xi, mu1, mu2, I0, d, phi;
    int32_t fric;
    int32_t nx, ny, ntx, nty, np;
    // driver state
    double time, dt;
    int32_t frame, step, epoch, parity;
    int32_t stop, out_step, keep_grid, frame_hit;
    int32_t steps_in_call, err, n_ptiles, n_gtiles;
    int32_t last_out_epoch, pad0;
    unsigned long long vmax2_bits;   // max |v_p|^2 of the positions to be used next (G2P output)
    unsigned long long bb_next[4];   // orderable bits of min x, max x, min y, max y (next positions)
    unsigned long long bb_cur[4];    // ... of the positions used by the last P2G
};

// Order-preserving map of a double to uint64 (for min/max with integer atomics).
__host__ __device__ inline unsigned long long ord_bits(double v) {
    union { double d; unsigned long long u; } c;
    c.d = v;
    return (c.u >> 63) ? ~c.u : (c.u | 0x8000000000000000ull);
}
__host__ __device__ inline double ord_value(unsigned long long k) {
    union { double d; unsigned long long u; } c;
    c.u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return c.d;
}

} // namespace slab
