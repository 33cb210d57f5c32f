// slab_kernels.hip -- gfx950 kernels of the depth-averaged MPM step, FP64 form.
//
// One step (src/simulation.cpp:524-563) is the launch sequence
//   ctl_begin -> tile_reduce -> tile_scan -> tile_offsets -> scatter
//   -> tile_sort -> p2g_grid -> g2p -> ctl_end
// Every kernel first reads Ctl::stop and returns when it is set, so a host can
// enqueue many steps without synchronising; a step that reaches an output
// frame (or total_time) stops the rest of the queue.
//
// The FP64 form restates the reference arithmetic operation by operation and
// is compiled with -ffp-contract=off: apart from the order in which P2G sums
// particle contributions into a node (here: by cell, then by creation index;
// in the reference: storage order), results match the CPU restatement bit for
// bit.  The summation order is canonical -- it depends only on positions and
// creation indices -- so results are reproducible run to run.
#include "slab_common.hpp"
#include "slab_kernels.hpp"
#include "slab_device.hpp"

#pragma clang fp contract(off)

namespace slab {

// ---------------------------------------------------------------- helpers

// Quadratic B-spline, src/tools.hpp:119-128.  Both pieces are evaluated with
// the reference's operations and the piece is selected (no branch).
__device__ __forceinline__ double Nq(double u) {
    const double a = fabs(u);
    const double inner = 0.75 - a * a;
    const double outer = 0.5 * (1.5 - a) * (1.5 - a);
    return a < 0.5 ? inner : (a < 1.5 ? outer : 0.0);
}
// Its derivative, src/tools.hpp:131-140 (sgn as src/tools.hpp:90-92).
__device__ __forceinline__ double dNq(double u) {
    const double a = fabs(u);
    const double inner = -2.0 * u;
    const double outer = u - 1.5 * (double)((0.0 < u) - (u < 0.0));
    return a < 0.5 ? inner : (a < 1.5 ? outer : 0.0);
}

// Stencil base trunc((x - x0)/dx - 0.5) of src/simulation.cpp:741-742 in
// terrain coordinates (x0 = 0), clamped so the 3x3 stencil stays on the
// terrain (the reference leaves out-of-terrain positions undefined).  NaN -> 0.
__device__ __forceinline__ int base_index(double x, double inv, int nmax) {
    double f = x * inv - 0.5;
    f = fmin(fmax(f, 0.0), (double)nmax);
    return (int)f;
}
// Floor cell of particleGradients / particleSlopeAngle,
// src/complex_terrain.cpp:113-130, clamped to the terrain.
__device__ __forceinline__ int floor_index(double x, double inv, int nmax) {
    double f = floor(x * inv);
    f = fmin(fmax(f, 0.0), (double)nmax);
    return (int)f;
}

__device__ __forceinline__ PBuf64 cur_buf(const Ctl* c, const PBuf64& A, const PBuf64& B) {
    return c->parity ? B : A;
}
__device__ __forceinline__ PBuf64 nxt_buf(const Ctl* c, const PBuf64& A, const PBuf64& B) {
    return c->parity ? A : B;
}

// Key of a position: cell = stencil base node, grouped by 16x16 tile.
__device__ __forceinline__ int cell_key(const Ctl* c, double x, double y, int* tile_out) {
    const int ib = base_index(x, c->inv_dx, c->nx - 3);
    const int jb = base_index(y, c->inv_dx, c->ny - 3);
    const int t = (ib >> 4) * c->nty + (jb >> 4);
    *tile_out = t;
    return t * TILE_CELLS + (((ib & 15) << 4) | (jb & 15));
}

// ---------------------------------------------------------------- control

__global__ void k_ctl_clear(Ctl* c) {
    if (c->stop == STOP_FRAME) c->stop = RUNNING;
    c->steps_in_call = 0;
    c->frame_hit = 0;
}

// updateDt, src/simulation.cpp:565-582, plus the bookkeeping of one step.
// One block of 256 threads: reduces the partial shards of the previous g2p (or
// of k_keys after upload) to max |v_p|^2 and the bounding box of the positions
// this step's P2G uses; a step that proceeds resets the shards for its g2p
// (a stopped one keeps them for diagnostics).
// Run by one workgroup: the extra block of k_tile_reduce (the tile counting
// does not read what this writes: its counts come from the previous g2p).
__device__ void ctl_begin_block(Ctl* c, unsigned long long* __restrict__ part) {
    __shared__ unsigned long long s_out[5];
    __shared__ int s_go;
    reduce_partials(part, s_out);
    if (threadIdx.x == 0) {
        s_go = 0;
        if (!(c->time < c->total_time)) { // while condition, simulation.cpp:500
            c->stop = STOP_DONE;
        } else if (!(__longlong_as_double(s_out[4]) <= 1.7976931348623157e308)) {
            c->vmax2_bits = s_out[4];
            c->err = 5; // SLAB_ENAN
            c->stop = STOP_ERROR;
        } else {
            s_go = 1;
        }
    }
    __syncthreads();
    if (!s_go) return;
    reset_partials(part);
    if (threadIdx.x != 0) return;
    c->vmax2_bits = s_out[4];
    const double v2 = __longlong_as_double(s_out[4]);
    const double max_speed = sqrt(v2);
    double dt;
    if (max_speed > 0.0) {
        const double a = c->C_cfl * c->dx / max_speed;
        dt = (c->dt_max < a) ? c->dt_max : a;
    } else {
        dt = c->dt_max;
    }
    const double next_frame_time = c->output_dt * (double)(c->frame + 1);
    if (c->time + dt > next_frame_time - 1e-12) dt = next_frame_time - c->time;
    c->dt = dt;
    c->step += 1;
    c->epoch += 1;
    const double tnew = c->time + dt;
    const int hit = (c->output_dt * (double)(c->frame + 1) - tnew < 1e-13);
    c->out_step = hit || c->keep_grid;
    if (c->out_step) c->last_out_epoch = c->epoch;
    for (int k = 0; k < 4; ++k) c->bb_cur[k] = s_out[k];
}

// time += dt and the frame trigger, src/simulation.cpp:503-509.
__global__ void k_ctl_end(Ctl* c) {
    if (c->stop) return;
    c->time += c->dt;
    c->parity ^= 1;
    c->steps_in_call += 1;
    if (c->output_dt * (double)(c->frame + 1) - c->time < 1e-13) {
        c->frame += 1;
        c->frame_hit = 1;
        c->stop = STOP_FRAME;
    }
}

__global__ void k_prime_pre(Ctl* c) { c->stop = RUNNING; }

// simulate() priming with dt = 0 (src/simulation.cpp:493-495, simulation.hpp:39).
__global__ void __launch_bounds__(WG) k_ctl_prime(Ctl* c, const unsigned long long* __restrict__ part) {
    __shared__ unsigned long long s_out[5];
    reduce_partials(part, s_out); // kept: step 1's ctl_begin reads them again
    if (threadIdx.x != 0) return;
    c->dt = 0.0;
    c->epoch += 1;
    c->out_step = 1;
    c->last_out_epoch = c->epoch;
    c->vmax2_bits = s_out[4];
    for (int k = 0; k < 4; ++k) c->bb_cur[k] = s_out[k];
}

// Keys, tile counts, |v|^2 max and bounding box of the current particles
// (used once after upload; afterwards g2p produces them).  Partial shards
// start at the identity (k_part_init).
__global__ void __launch_bounds__(WG) k_keys(Ctl* c, PBuf64 A, PBuf64 B, int* __restrict__ key,
                                             int* __restrict__ tile_count, unsigned long long* __restrict__ part) {
    const PBuf64 P = cur_buf(c, A, B);
    const int s = blockIdx.x * WG + threadIdx.x;
    const bool valid = s < c->np;
    double v2 = 0.0;
    unsigned long long bx0 = ~0ull, bx1 = 0, by0 = ~0ull, by1 = 0;
    int t = 0, k = 0;
    if (valid) {
        const double x = P.x[s], y = P.y[s];
        k = cell_key(c, x, y, &t);
        const double vx = P.vx[s], vy = P.vy[s];
        v2 = vx * vx + vy * vy;
        bx0 = bx1 = ord_bits(x);
        by0 = by1 = ord_bits(y);
    }
    wave_tile_count(t, valid, tile_count);
    if (valid) key[s] = k;
    block_partials(v2, bx0, bx1, by0, by1, part);
}

__global__ void k_part_init(unsigned long long* __restrict__ part) { reset_partials(part); }

// Initial particle z: 4x4 quadratic interpolation of terrain z,
// src/simulation.cpp:455-465.
__global__ void k_init_z(Ctl* c, PBuf64 A, const double* __restrict__ tz) {
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= c->np) return;
    const double x = A.x[s], y = A.y[s], inv = c->inv_dx, dx = c->dx;
    const int ib = (int)((x - 0.0) * inv) - 1;
    const int jb = (int)((y - 0.0) * inv) - 1;
    double z = 0.0;
    for (int i = ib; i < ib + 4; ++i) {
        for (int j = jb; j < jb + 4; ++j) {
            if (i < 0 || j < 0 || i >= c->nx || j >= c->ny) continue;
            const double w = Nq((x - (double)i * dx) * inv) * Nq((y - (double)j * dx) * inv);
            z += tz[(size_t)i * c->ny + j] * w;
        }
    }
    A.z[s] = z;
}

// ---------------------------------------------------------------- tile scan

// Block-wide exclusive scan of 3 channels over 256 threads (4 waves).
__device__ __forceinline__ void block_exscan3(int v[3], int out[3], int tot[3]) {
    __shared__ int s_w[4][3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc[3] = {v[0], v[1], v[2]};
    for (int o = 1; o < 64; o <<= 1) {
        for (int k = 0; k < 3; ++k) {
            int u = __shfl_up(inc[k], o, 64);
            if (lane >= o) inc[k] += u;
        }
    }
    if (lane == 63)
        for (int k = 0; k < 3; ++k) s_w[wv][k] = inc[k];
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        int pre = 0, all = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wv) pre += s_w[w][k];
            all += s_w[w][k];
        }
        out[k] = pre + inc[k] - v[k];
        tot[k] = all;
    }
    __syncthreads();
}

__device__ __forceinline__ void tile_flags(const Ctl* c, const int* __restrict__ tile_count, int t, int ntiles,
                                           int f[3]) {
    f[0] = f[1] = f[2] = 0;
    if (t >= ntiles) return;
    const int nty = c->nty;
    const int tx = t / nty, ty = t - tx * nty;
    const int n = tile_count[t];
    int g = n > 0;
    if (tx > 0) g |= tile_count[t - nty] > 0;
    if (ty > 0) g |= tile_count[t - 1] > 0;
    if (tx > 0 && ty > 0) g |= tile_count[t - nty - 1] > 0;
    f[0] = n;
    f[1] = n > 0;
    f[2] = g;
}

// Per-block sums of (particles, particle tiles, grid tiles).  With
// with_begin != 0 the last block starts the step instead (ctl_begin_block).
__global__ void __launch_bounds__(WG) k_tile_reduce(Ctl* c, const int* __restrict__ tile_count, int ntiles,
                                                    int* __restrict__ bsum, unsigned long long* __restrict__ part,
                                                    int with_begin) {
    if (c->stop) return;
    if (with_begin && blockIdx.x == gridDim.x - 1) {
        ctl_begin_block(c, part);
        return;
    }
    int acc[3] = {0, 0, 0};
    const int t0 = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int f[3];
        tile_flags(c, tile_count, t0 + k, ntiles, f);
        acc[0] += f[0];
        acc[1] += f[1];
        acc[2] += f[2];
    }
    int ex[3], tot[3];
    block_exscan3(acc, ex, tot);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) bsum[3 * blockIdx.x + k] = tot[k];
}

// Exclusive scan of the block sums (one workgroup, any number of blocks).
__global__ void k_tile_scan(Ctl* c, const int* __restrict__ bsum, int nblk, int* __restrict__ boff) {
    if (c->stop) return;
    int run[3] = {0, 0, 0};
    for (int b0 = 0; b0 < nblk; b0 += WG) {
        const int b = b0 + threadIdx.x;
        int v[3] = {0, 0, 0};
        if (b < nblk)
            for (int k = 0; k < 3; ++k) v[k] = bsum[3 * b + k];
        int ex[3], tot[3];
        block_exscan3(v, ex, tot);
        if (b < nblk)
            for (int k = 0; k < 3; ++k) boff[3 * b + k] = run[k] + ex[k];
        for (int k = 0; k < 3; ++k) run[k] += tot[k];
    }
    if (threadIdx.x == 0) {
        c->n_ptiles = run[1];
        c->n_gtiles = run[2];
    }
}

// tile_begin, particle-tile list, grid-tile list (both in tile-id order) and
// stamps.  The block's starting offsets come from k_tile_scan (boff), or, with
// few scan blocks (merged != 0), from block_prefix3, block 0 then writing the
// list lengths.
__global__ void __launch_bounds__(WG) k_tile_offsets(Ctl* c, const int* __restrict__ tile_count, int ntiles,
                                                     const int* __restrict__ boff, const int* __restrict__ bsum,
                                                     int nblk, int merged, int* __restrict__ tile_begin,
                                                     int* __restrict__ ptiles, int* __restrict__ gtiles,
                                                     int* __restrict__ pstamp, int* __restrict__ gstamp,
                                                     int* __restrict__ tile_pos) {
    if (c->stop) return;
    int base[3];
    if (merged) {
        int tot[3];
        block_prefix3(bsum, nblk, base, tot);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->n_ptiles = tot[1];
            c->n_gtiles = tot[2];
        }
    } else {
        for (int q = 0; q < 3; ++q) base[q] = boff[3 * blockIdx.x + q];
    }
    const int t0 = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    int f[SCAN_ITEMS][3];
    int acc[3] = {0, 0, 0};
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        tile_flags(c, tile_count, t0 + k, ntiles, f[k]);
        for (int q = 0; q < 3; ++q) acc[q] += f[k][q];
    }
    int ex[3], tot[3];
    block_exscan3(acc, ex, tot);
    int run[3];
    for (int q = 0; q < 3; ++q) run[q] = base[q] + ex[q];
    const int ep = c->epoch;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int t = t0 + k;
        if (t >= ntiles) break;
        tile_begin[t] = run[0];
        if (f[k][1]) {
            ptiles[run[1]] = t;
            tile_pos[t] = run[1];
            pstamp[t] = ep;
        }
        if (f[k][2]) {
            gtiles[run[2]] = t;
            gstamp[t] = ep;
        }
        for (int q = 0; q < 3; ++q) run[q] += f[k][q];
    }
}

// ---------------------------------------------------------------- sort

// Buckets the storage slots by tile: each particle takes the next slot of its
// tile from a per-tile cursor (the tile sort clears the cursors of the tiles
// it orders).  A wave handles SC_PER groups of 64 consecutive slots and issues
// the cursor atomics of all groups (one per run of equal tiles) before it
// waits for any of them.  Arrival order inside a tile is arbitrary: the tile
// sort makes it canonical.
constexpr int SC_PER = kScatterPerThread;
__global__ void __launch_bounds__(WG) k_scatter(Ctl* c, const int* __restrict__ key,
                                                const int* __restrict__ tile_begin, int* __restrict__ cursor,
                                                int* __restrict__ tmp) {
    if (c->stop) return;
    const int n = c->np, lane = threadIdx.x & 63;
    const int s0 = (blockIdx.x * WG + (threadIdx.x & ~63)) * SC_PER + lane;
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    int t[SC_PER], start[SC_PER], base[SC_PER];
    bool valid[SC_PER];
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) {
        const int s = s0 + 64 * j;
        valid[j] = s < n;
        t[j] = valid[j] ? key[s] >> 8 : 0;
    }
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) {
        const int tprev = __shfl_up(t[j], 1, 64);
        const bool head = valid[j] && (lane == 0 || tprev != t[j]);
        const unsigned long long vmask = __ballot(valid[j]);
        const unsigned long long hmask = __ballot(head);
        const unsigned long long hle = hmask & upto;
        const unsigned long long brk = (hmask | ~vmask) & ~upto; // next head or first invalid lane
        start[j] = hle ? 63 - __clzll((long long)hle) : 0;
        const int end = brk ? (__ffsll((long long)brk) - 1) : 64;
        base[j] = head ? atomicAdd(&cursor[t[j]], end - lane) : 0;
    }
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) {
        const int b = __shfl(base[j], start[j], 64);
        if (valid[j]) tmp[tile_begin[t[j]] + b + (lane - start[j])] = s0 + 64 * j;
    }
}

// Per particle tile: counting sort by cell in LDS, then each cell's run sorted
// by creation index, giving the canonical order (tile, cell, creation index).
// Writes idx (sorted position -> storage slot) and cell_begin, and clears the
// tile's count (for the next step's g2p) and its scatter cursor.
// Tiles of at most TS_CAP particles are staged in LDS (cell, creation index,
// storage index), counted and placed by cell there, and each particle is
// written at its cell's start plus its rank by creation index in the cell;
// larger tiles place by cell and insertion-sort each cell in global memory.  The dependent global reads (bucket, then key and creation
// index) are issued in groups of TS_GROUP per thread.
constexpr int TS_CAP = 2048;
constexpr int TS_GROUP = 4;

__global__ void __launch_bounds__(WG) k_tile_sort(Ctl* c, const int32_t* __restrict__ pidA,
                                                  const int32_t* __restrict__ pidB, const int* __restrict__ key,
                                                  const int* __restrict__ ptiles, int* __restrict__ tile_count,
                                                  int* __restrict__ cursor, const int* __restrict__ tile_begin,
                                                  const int* __restrict__ tmp, int* __restrict__ idx,
                                                  int* __restrict__ cell_begin) {
    if (c->stop) return;
    __shared__ int s_hist[TILE_CELLS];
    __shared__ int s_off[TILE_CELLS];
    __shared__ int s_n;
    __shared__ int s_pid[TS_CAP];
    __shared__ int s_sv[TS_CAP];
    __shared__ unsigned short s_ord[TS_CAP];
    __shared__ unsigned char s_cell[TS_CAP];
    const int32_t* __restrict__ pid = c->parity ? pidB : pidA;
    const int nt = c->n_ptiles;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = ptiles[w];
        if (threadIdx.x == 0) s_n = tile_count[t];
        s_hist[threadIdx.x] = 0;
        __syncthreads();
        const int n = s_n;
        const int b = tile_begin[t];
        const bool small = n <= TS_CAP;
        for (int r0 = threadIdx.x; r0 < n; r0 += WG * TS_GROUP) {
            int sv[TS_GROUP], kc[TS_GROUP], pd[TS_GROUP];
#pragma unroll
            for (int j = 0; j < TS_GROUP; ++j) sv[j] = tmp[b + min(r0 + j * WG, n - 1)];
#pragma unroll
            for (int j = 0; j < TS_GROUP; ++j) {
                kc[j] = key[sv[j]] & 255;
                pd[j] = small ? pid[sv[j]] : 0;
            }
#pragma unroll
            for (int j = 0; j < TS_GROUP; ++j) {
                const int r = r0 + j * WG;
                if (r >= n) continue;
                atomicAdd(&s_hist[kc[j]], 1);
                if (small) {
                    s_cell[r] = (unsigned char)kc[j];
                    s_pid[r] = pd[j];
                    s_sv[r] = sv[j];
                }
            }
        }
        __syncthreads();
        // exclusive scan of the 256 cell counts
        {
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            const int v = s_hist[threadIdx.x];
            int inc = v;
            for (int o = 1; o < 64; o <<= 1) {
                int u = __shfl_up(inc, o, 64);
                if (lane >= o) inc += u;
            }
            __shared__ int s_w[4];
            if (lane == 63) s_w[wv] = inc;
            __syncthreads();
            int pre = 0;
            for (int q = 0; q < wv; ++q) pre += s_w[q];
            s_off[threadIdx.x] = pre + inc - v;
            cell_begin[(size_t)t * CELL_STRIDE + threadIdx.x] = b + pre + inc - v;
            if (threadIdx.x == 0) {
                cell_begin[(size_t)t * CELL_STRIDE + TILE_CELLS] = b + n;
                tile_count[t] = 0;
                cursor[t] = 0;
            }
        }
        __syncthreads();
        // s_hist becomes the per-cell cursor
        s_hist[threadIdx.x] = s_off[threadIdx.x];
        __syncthreads();
        if (small) {
            for (int r = threadIdx.x; r < n; r += WG) s_ord[atomicAdd(&s_hist[s_cell[r]], 1)] = (unsigned short)r;
            __syncthreads();
            // s_off[c] .. s_hist[c] lists cell c's arrivals; each arrival goes
            // to the cell start plus its rank there by creation index (unique),
            // counted by its own thread: no cell is sorted serially
            for (int r = threadIdx.x; r < n; r += WG) {
                const int cr = s_cell[r], pr = s_pid[r];
                const int lo = s_off[cr], hi = s_hist[cr];
                int rank = 0;
                for (int q = lo; q < hi; ++q) rank += s_pid[s_ord[q]] < pr ? 1 : 0;
                idx[b + lo + rank] = s_sv[r];
            }
        } else {
            for (int r0 = threadIdx.x; r0 < n; r0 += WG * TS_GROUP) {
                int sv[TS_GROUP], kc[TS_GROUP];
#pragma unroll
                for (int j = 0; j < TS_GROUP; ++j) sv[j] = tmp[b + min(r0 + j * WG, n - 1)];
#pragma unroll
                for (int j = 0; j < TS_GROUP; ++j) kc[j] = key[sv[j]] & 255;
#pragma unroll
                for (int j = 0; j < TS_GROUP; ++j)
                    if (r0 + j * WG < n) idx[b + atomicAdd(&s_hist[kc[j]], 1)] = sv[j];
            }
            __syncthreads();
            // insertion sort of each cell's run by creation index
            const int lo = b + s_off[threadIdx.x];
            const int hi = b + s_hist[threadIdx.x];
            for (int q = lo + 1; q < hi; ++q) {
                const int sv = idx[q];
                const int pv = pid[sv];
                int k = q - 1;
                while (k >= lo && pid[idx[k]] > pv) {
                    idx[k + 1] = idx[k];
                    --k;
                }
                idx[k + 1] = sv;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- P2G + grid

// P2G (src/simulation.cpp:726-798) per particle tile, without atomics: one
// thread per stencil-base cell sums its particles (canonical order, read
// through idx) into its 9 stencil slots in registers; 9 phases then add slot
// (d,e) of every cell into node cell+(d,e) of an 18x18-node LDS region (one-to-
// one within a phase, fixed order) and the region is written as the tile's
// partial.  The 7 FP64 fields run in two groups (mass + momenta, then height +
// forces) to bound the registers; the per-particle terms of
// src/simulation.cpp:732-753 are formed in the group that uses them, with the
// same operations (so the same values) as a separate per-particle pass would.
// k_grid64 merges the partials per node.
constexpr int P2G64_FIELDS = 7;
constexpr int REGION64 = TILE + 2;
constexpr int REGION64_NODES = REGION64 * REGION64;

__global__ void __launch_bounds__(WG) k_p2g64(Ctl* c, PBuf64 A, PBuf64 B, const int* __restrict__ idx,
                                              const double* __restrict__ tgx, const double* __restrict__ tgy,
                                              const double* __restrict__ tcth, const int* __restrict__ ptiles,
                                              const int* __restrict__ cell_begin, double* __restrict__ tpart) {
    if (c->stop) return;
    __shared__ double s_acc[4 * REGION64_NODES];
    const PBuf64 P = cur_buf(c, A, B);
    const int nt = c->n_ptiles, nty = c->nty, nx = c->nx, ny = c->ny;
    const double inv = c->inv_dx, dx = c->dx, g = c->gravity, rho = c->rho;
    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = ptiles[w];
        const int tx = t / nty, ty = t - tx * nty;
        const int ib = tx * TILE + lx, jb = ty * TILE + ly;
        const int* cb = cell_begin + (size_t)t * CELL_STRIDE; // indexed by tile id (k_tile_sort)
        const int r0 = cb[threadIdx.x], r1 = cb[threadIdx.x + 1];
        double* dst = tpart + (size_t)w * (P2G64_FIELDS * REGION64_NODES);
        double X[3], Yn[3];
        for (int d = 0; d < 3; ++d) {
            X[d] = (double)(ib + d) * dx;
            Yn[d] = (double)(jb + d) * dx;
        }
        for (int grp = 0; grp < 2; ++grp) {
            const int nf = grp == 0 ? 4 : 3;
            for (int q = threadIdx.x; q < 4 * REGION64_NODES; q += WG) s_acc[q] = 0.0;
            double acc[9][4];
#pragma unroll
            for (int k = 0; k < 9; ++k)
#pragma unroll
                for (int f = 0; f < 4; ++f) acc[k][f] = 0.0;
            for (int r = r0; r < r1; ++r) {
                const int sn = idx[r];
                const double xp = P.x[sn], yp = P.y[sn];
                double f0, f1, f2, f3 = 0.0;
                if (grp == 0) {
                    f0 = P.m[sn]; f1 = P.vx[sn]; f2 = P.vy[sn]; f3 = P.vz[sn];
                } else {
                    f0 = P.h[sn]; f1 = P.A[sn]; f2 = P.m[sn];
                }
                double ux[3], uy[3], wx[3], wy[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    ux[d] = (xp - X[d]) * inv;
                    uy[d] = (yp - Yn[d]) * inv;
                    wx[d] = Nq(ux[d]);
                    wy[d] = Nq(uy[d]);
                }
                if (grp == 0) {
                    const double m = f0, vmx = f1 * f0, vmy = f2 * f0, vzm = f3 * f0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            const double wgt = wx[d] * wy[e];
                            if (!(wgt > 3e-3)) continue; // simulation.cpp:780
                            double* a = acc[d * 3 + e];
                            a[0] += m * wgt;
                            a[1] += vmx * wgt;
                            a[2] += vmy * wgt;
                            a[3] += vzm * wgt;
                        }
                    }
                } else {
                    const double hp = f0, Ap = f1, mp = f2;
                    // terrain of the particle's floor cell (src/complex_terrain.cpp:113-130)
                    const size_t fc = (size_t)floor_index(xp, inv, nx - 1) * ny + floor_index(yp, inv, ny - 1);
                    const double cth = tcth[fc];
                    const double coeff = -mp * g * cth * cth;
                    const double hm = hp * mp, hA = hp * Ap, mbx = coeff * tgx[fc], mby = coeff * tgy[fc];
                    const double s00 = -0.5 * rho * g * cth * cth * hp;
                    double dwx[3], dwy[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        dwx[d] = dNq(ux[d]) * inv;
                        dwy[d] = dNq(uy[d]) * inv;
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            const double wgt = wx[d] * wy[e];
                            if (!(wgt > 3e-3)) continue;
                            const double gwx = dwx[d] * wy[e], gwy = dwy[e] * wx[d];
                            double* a = acc[d * 3 + e];
                            a[0] += hm * wgt;
                            a[1] += mbx * wgt - hA * (s00 * gwx + 0.0 * gwy);
                            a[2] += mby * wgt - hA * (s00 * gwy + 0.0 * gwx);
                        }
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int ph = 0; ph < 9; ++ph) {
                const int node = (lx + ph / 3) * REGION64 + (ly + ph % 3);
#pragma unroll
                for (int f = 0; f < 4; ++f)
                    if (f < nf) s_acc[f * REGION64_NODES + node] += acc[ph][f];
                __syncthreads();
            }
            double* gd = dst + (size_t)(grp == 0 ? 0 : 4) * REGION64_NODES;
            for (int q = threadIdx.x; q < nf * REGION64_NODES; q += WG) gd[q] = s_acc[q];
            __syncthreads();
        }
    }
}

// computeGridVelocity (src/simulation.cpp:800-875) per grid tile, one thread
// per node, after merging the node's P2G sums from the particle tiles that
// reach it (own, left, below, diagonal; fixed order).
__global__ void __launch_bounds__(WG) k_grid64(Ctl* c, const int* __restrict__ gtiles, const int* __restrict__ pstamp,
                                               const int* __restrict__ tile_pos, const double* __restrict__ tpart,
                                               const double* __restrict__ tgx, const double* __restrict__ tgy,
                                               const double* __restrict__ tcth, double* __restrict__ h_max,
                                               double* __restrict__ v_max2, double4* __restrict__ vf,
                                               double* __restrict__ out_m, double* __restrict__ out_h,
                                               double* __restrict__ out_vz) {
    if (c->stop) return;
    const int nt = c->n_gtiles, ep = c->epoch, nx = c->nx, ny = c->ny, nty = c->nty;
    const double g = c->gravity, dt = c->dt;
    const int out = c->out_step;
    const int a = threadIdx.x >> 4, b = threadIdx.x & 15;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = gtiles[w];
        const int tx = t / nty, ty = t - tx * nty;
        const int i = tx * TILE + a, j = ty * TILE + b;
        if (i >= nx || j >= ny) continue;
        double acc[P2G64_FIELDS] = {0, 0, 0, 0, 0, 0, 0};
        int src[4], rn[4];
        int ns = 0;
        src[ns] = t;
        rn[ns++] = a * REGION64 + b;
        if (tx > 0 && a < 2) {
            src[ns] = t - nty;
            rn[ns++] = (a + TILE) * REGION64 + b;
        }
        if (ty > 0 && b < 2) {
            src[ns] = t - 1;
            rn[ns++] = a * REGION64 + (b + TILE);
        }
        if (tx > 0 && ty > 0 && a < 2 && b < 2) {
            src[ns] = t - nty - 1;
            rn[ns++] = (a + TILE) * REGION64 + (b + TILE);
        }
        for (int k = 0; k < ns; ++k) {
            if (pstamp[src[k]] != ep) continue;
            const double* q = tpart + (size_t)tile_pos[src[k]] * (P2G64_FIELDS * REGION64_NODES) + rn[k];
#pragma unroll
            for (int f = 0; f < P2G64_FIELDS; ++f) acc[f] += q[f * REGION64_NODES];
        }
        // acc: m, P_x, P_y, P_z | H, F_x, F_y
        const double am = acc[0], apx = acc[1], apy = acc[2], apz = acc[3], ah = acc[4], afx = acc[5], afy = acc[6];
        const size_t k = (size_t)i * ny + j;
        double nvx = 0.0, nvy = 0.0, flx = 0.0, fly = 0.0, hh = 0.0, vz = 0.0;
        if (am > 1e-6) {
            const double vx = apx / am, vy = apy / am;
            vz = apz / am;
            hh = ah / am;
            const double ax = afx / am, ay = afy / am;
            const double vtx = vx + dt * ax, vty = vy + dt * ay;
            const double vsq = vtx * vtx + vty * vty;
            if (vsq > 1e-16) {
                const double Gx = tgx[k], Gy = tgy[k], cth = tcth[k];
                const double v_z = vtx * Gx + vty * Gy;
                const double v3sq = vsq + v_z * v_z;
                const double vn = sqrt(v3sq);
                double af;
                if (c->fric == 0) {
                    af = c->mu * g * cth;
                } else if (c->fric == 1) {
                    af = c->mu * g * cth + g * v3sq / (c->xi * hh * cth);
                } else {
                    const double I = 5. * c->d * vn / (2. * hh * cth * cth * sqrt(g * hh * c->phi));
                    const double muI = c->mu1 + (c->mu2 - c->mu1) / (1. + c->I0 / I);
                    af = muI * g * cth;
                }
                const double dv = dt * af;
                if (dv < vn) {
                    const double f = 1. - dv / vn;
                    nvx = vtx * f;
                    nvy = vty * f;
                }
            }
            flx = nvx - vx;
            fly = nvy - vy;
            const double hm0 = h_max[k];
            if (hm0 < hh) h_max[k] = hh;
            const double v2 = nvx * nvx + nvy * nvy;
            const double vm0 = v_max2[k];
            if (vm0 < v2) v_max2[k] = v2;
        }
        vf[k] = make_double4(nvx, nvy, flx, fly);
        if (out) {
            out_m[k] = am;
            out_h[k] = hh;
            out_vz[k] = vz;
        }
    }
}

// One thread per particle in sorted order (src/simulation.cpp:877-940).  Reads
// the particle from the current buffer through idx and writes the updated
// particle to slot r of the other buffer, so storage follows the sort.  Also
// produces the next step's keys / tile counts, max |v|^2 and bounding box.
__global__ void __launch_bounds__(WG) k_g2p(Ctl* c, PBuf64 A, PBuf64 B, const int* __restrict__ idx,
                                            const double4* __restrict__ vf, const double* __restrict__ tz,
                                            const double* __restrict__ tgx, const double* __restrict__ tgy,
                                            int* __restrict__ key, int* __restrict__ tile_count,
                                            unsigned long long* __restrict__ part) {
    if (c->stop) return;
    const PBuf64 P = cur_buf(c, A, B);
    const PBuf64 Q = nxt_buf(c, A, B);
    const int r = blockIdx.x * WG + threadIdx.x;
    const bool valid = r < c->np;
    double v2 = 0.0;
    unsigned long long bx0 = ~0ull, bx1 = 0, by0 = ~0ull, by1 = 0;
    int t = 0, kk = 0;
    if (valid) {
        const int nx = c->nx, ny = c->ny;
        const double inv = c->inv_dx, dx = c->dx, dt = c->dt, fr = c->flip_ratio;
        const int s = idx[r];
        const double xp = P.x[s], yp = P.y[s];
        const size_t fc = (size_t)floor_index(xp, inv, nx - 1) * ny + floor_index(yp, inv, ny - 1);
        const double Gx = tgx[fc], Gy = tgy[fc];
        const int ib = base_index(xp, inv, nx - 3), jb = base_index(yp, inv, ny - 3);
        double wx[3], wy[3], dwx[3], dwy[3];
        for (int d = 0; d < 3; ++d) {
            const double u = (xp - (double)(ib + d) * dx) * inv;
            wx[d] = Nq(u);
            dwx[d] = dNq(u) * inv;
            const double v = (yp - (double)(jb + d) * dx) * inv;
            wy[d] = Nq(v);
            dwy[d] = dNq(v) * inv;
        }
        double picx = 0.0, picy = 0.0, flx = 0.0, fly = 0.0, zp = 0.0, tr = 0.0;
        for (int di = 0; di < 3; ++di) {
            for (int dj = 0; dj < 3; ++dj) {
                const size_t k = (size_t)(ib + di) * ny + (jb + dj);
                const double w = wx[di] * wy[dj];
                const double gwx = dwx[di] * wy[dj];
                const double gwy = dwy[dj] * wx[di];
                const double4 q = vf[k];
                picx += q.x * w;
                picy += q.y * w;
                flx += q.z * w;
                fly += q.w * w;
                zp += tz[k] * w;
                tr += gwx * q.x + gwy * q.y;
            }
            // one stencil row's gathers in flight at a time: 104 instead of 132
            // registers, 4 waves per SIMD instead of 3 (same arithmetic)
            __builtin_amdgcn_sched_barrier(0);
        }
        tr *= dt;
        const double nvx = fr * (P.vx[s] + flx) + (1. - fr) * picx;
        const double nvy = fr * (P.vy[s] + fly) + (1. - fr) * picy;
        const double xn = xp + dt * picx, yn = yp + dt * picy;
        const double hn = P.h[s] / (1. + tr);
        const double mp = P.m[s];
        Q.x[r] = xn;
        Q.y[r] = yn;
        Q.vx[r] = nvx;
        Q.vy[r] = nvy;
        Q.vz[r] = nvx * Gx + nvy * Gy;
        Q.z[r] = zp;
        Q.h[r] = hn;
        Q.A[r] = mp / (hn * c->rho);
        Q.m[r] = mp;
        Q.pid[r] = P.pid[s];
        kk = cell_key(c, xn, yn, &t);
        v2 = nvx * nvx + nvy * nvy;
        bx0 = bx1 = ord_bits(xn);
        by0 = by1 = ord_bits(yn);
    }
    wave_tile_count(t, valid, tile_count);
    if (valid) key[r] = kk;
    block_partials(v2, bx0, bx1, by0, by1, part);
}

// ---------------------------------------------------------------- downloads / diagnostics

// Gather the reference's remesh window (src/simulation.cpp:594-612) out of the
// full-terrain node arrays; nodes outside this step's grid tiles are empty.
__global__ void k_window(const Ctl* c, int i0, int j0, int wnx, int wny, const int* __restrict__ gstamp,
                         const double4* __restrict__ vf, const double* __restrict__ gm,
                         const double* __restrict__ gh, const double* __restrict__ gvz, double* __restrict__ o_h,
                         double* __restrict__ o_vx, double* __restrict__ o_vy, double* __restrict__ o_vz,
                         double* __restrict__ o_m) {
    const int q = blockIdx.x * WG + threadIdx.x;
    if (q >= wnx * wny) return;
    const int a = q / wny, b = q - a * wny;
    const int i = i0 + a, j = j0 + b;
    const int t = (i >> 4) * c->nty + (j >> 4);
    double h = 0, vx = 0, vy = 0, vz = 0, m = 0;
    if (gstamp[t] == c->epoch) {
        const size_t k = (size_t)i * c->ny + j;
        const double4 v = vf[k];
        vx = v.x;
        vy = v.y;
        h = gh[k];
        vz = gvz[k];
        m = gm[k];
    }
    o_h[q] = h;
    o_vx[q] = vx;
    o_vy[q] = vy;
    o_vz[q] = vz;
    o_m[q] = m;
}

// FP64 conservation sums, deterministic: per-workgroup partials in a fixed
// tile assignment, summed in order by k_diag_final.
__global__ void __launch_bounds__(WG) k_diag_grid(const Ctl* c, const int* __restrict__ gtiles,
                                                  const double4* __restrict__ vf, const double* __restrict__ gm,
                                                  double* __restrict__ part) {
    __shared__ double s[4][WG];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const int nt = c->n_gtiles, nty = c->nty;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = gtiles[w];
        const int tx = t / nty, ty = t - tx * nty;
        const int i = tx * TILE + (threadIdx.x >> 4), j = ty * TILE + (threadIdx.x & 15);
        if (i < c->nx && j < c->ny) {
            const size_t k = (size_t)i * c->ny + j;
            const double m = gm[k];
            const double4 v = vf[k];
            a0 += m;
            if (m > 1e-6) {
                a1 += m * v.x;
                a2 += m * v.y;
                a3 += 1.0;
            }
        }
    }
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    s[2][threadIdx.x] = a2;
    s[3][threadIdx.x] = a3;
    __syncthreads();
    for (int o = WG / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) part[4 * blockIdx.x + q] = s[q][0];
}

__global__ void __launch_bounds__(WG) k_diag_particles(const Ctl* c, PBuf64 A, PBuf64 B, double* __restrict__ part) {
    __shared__ double s[2][WG];
    const PBuf64 P = cur_buf(c, A, B);
    double a0 = 0, a1 = 0;
    for (int r = blockIdx.x * WG + threadIdx.x; r < c->np; r += gridDim.x * WG) {
        a0 += P.m[r];
        a1 += P.h[r] * P.A[r];
    }
    s[0][threadIdx.x] = a0;
    s[1][threadIdx.x] = a1;
    __syncthreads();
    for (int o = WG / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int q = 0; q < 2; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 2; ++q) part[2 * blockIdx.x + q] = s[q][0];
}

__global__ void __launch_bounds__(WG) k_diag_final(const Ctl* c, const double* __restrict__ pg, int ng,
                                                   const double* __restrict__ pp, int npb,
                                                   const unsigned long long* __restrict__ part, double* __restrict__ out) {
    __shared__ unsigned long long s_out[5];
    reduce_partials(part, s_out);
    if (threadIdx.x != 0) return;
    out[6] = sqrt(__longlong_as_double(s_out[4]));
    double g[4] = {0, 0, 0, 0}, p[2] = {0, 0};
    for (int b = 0; b < ng; ++b)
        for (int q = 0; q < 4; ++q) g[q] += pg[4 * b + q];
    for (int b = 0; b < npb; ++b)
        for (int q = 0; q < 2; ++q) p[q] += pp[2 * b + q];
    out[0] = g[0];
    out[1] = g[1];
    out[2] = g[2];
    out[3] = p[0];
    out[4] = p[1];
    out[7] = g[3];
}

} // namespace slab
