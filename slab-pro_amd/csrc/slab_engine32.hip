// slab_engine32.hip -- FP32 step engine for one or many scenarios on a shared
// DEM (single FP32 runs, the batched ensemble and the request server).
//
// Step launch sequence (fixed grid sizes; every kernel reads its counts from
// device memory, so a step can be replayed from a captured hipGraph):
//   begin -> tile_reduce -> tile_scan -> tile_offsets -> scatter -> tile_sort
//   -> p2g -> grid -> g2p -> end
// Positions are a packed floor cell (ci << 16 | cj) plus FP32 offsets in cell
// units; h*A = m / rho (A is not stored); time and dt are FP64 per slot.
// Multiply-add contraction is allowed in this translation unit.
#include "slab_engine32.hpp"
#include "slab_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <type_traits>

namespace slab {

// ============================================================== device code

__device__ __forceinline__ float Nf(float u) {
    const float a = fabsf(u);
    if (a < 0.5f) return 0.75f - a * a;
    if (a < 1.5f) return 0.5f * (1.5f - a) * (1.5f - a);
    return 0.0f;
}
__device__ __forceinline__ float dNf(float u) {
    const float a = fabsf(u);
    if (a < 0.5f) return -2.0f * u;
    if (a < 1.5f) return u - 1.5f * (float)((0.0f < u) - (u < 0.0f));
    return 0.0f;
}
// Weights / derivatives of the 3 stencil nodes of one axis for a position f
// relative to the base node (u = f - d), src/tools.hpp:119-140.  An unclamped
// base puts f in [0.5, 1.5], where u = f, f - 1, f - 2 fall in fixed pieces of
// N: the pieces are evaluated without branches, with the same operations as
// Nf / dNf (so the same results).
__device__ __forceinline__ void weights3f(float f, float inv, float w[3], float dw[3]) {
    const float t0 = 1.5f - f;        // 1.5 - |u0|
    const float u1 = f - 1.0f;        // |u1| <= 0.5
    const float t2 = 1.5f + (f - 2.0f); // 1.5 - |u2|
    w[0] = 0.5f * t0 * t0;
    w[1] = 0.75f - u1 * u1;
    w[2] = 0.5f * t2 * t2;
    dw[0] = (f - 1.5f) * inv;
    dw[1] = (-2.0f * u1) * inv;
    dw[2] = t2 * inv;
}

__device__ __forceinline__ PB32 ecur(const ECtl* c, const PB32& A, const PB32& B) { return c->parity ? B : A; }
__device__ __forceinline__ PB32 enxt(const ECtl* c, const PB32& A, const PB32& B) { return c->parity ? A : B; }

// Stencil base trunc(x/dx - 0.5) = ci - (o < 0.5) of src/simulation.cpp:741-742,
// clamped so the stencil stays on the terrain.
__device__ __forceinline__ int ebase(int ci, float o, int nmax) {
    int b = ci - (o < 0.5f ? 1 : 0);
    return b < 0 ? 0 : (b > nmax ? nmax : b);
}

__device__ __forceinline__ bool slot_active(int st) { return st == SLOT_PRIMING || st == SLOT_RUNNING; }

// The 8 XCDs (each with its own L2) take workgroups round-robin: block b runs
// on XCD b % 8.  G2P renumbers its blocks so that each XCD walks one
// contiguous eighth of the particles, whose grid nodes then stay in that XCD's
// L2.  (The scatter keeps the plain order: there the renumbering puts
// concurrent blocks on the same tile cursors and the atomics contend.)  The
// grid size is a multiple of 8.
constexpr int N_XCD = 8;
// nblk = blocks that hold work (the live count, not the capacity the grid is
// sized for): each XCD takes ceil(nblk / 8) consecutive ones; blocks past
// that get an index >= nblk (no work).
__device__ __forceinline__ int xcd_block(int nblk) {
    const int per = (nblk + N_XCD - 1) / N_XCD, j = (int)(blockIdx.x / N_XCD);
    return j < per ? (int)(blockIdx.x % N_XCD) * per + j : nblk;
}


// Element i of an array.  With O32 the caller guarantees i * sizeof(T) <
// 2^32: the access then takes the array base from scalar registers and a
// 32-bit byte offset, instead of 64-bit address arithmetic per lane.
template <bool O32, typename T>
__device__ __forceinline__ T& el(T* base, size_t i) {
    if constexpr (O32) {
        using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
        return *reinterpret_cast<T*>(reinterpret_cast<B*>(base) + (unsigned)i * (unsigned)sizeof(T));
    } else {
        return base[i];
    }
}

// The terrain and slot sizes keys are formed from, read from the control
// block once, before a kernel's first store (after one, the compiler can no
// longer read the block through the scalar cache and waits on the stores).
struct EGeo {
    int nx, ny, nty, ntd, nslots;
};
__device__ __forceinline__ EGeo egeo(const ECtl* c) { return {c->nx, c->ny, c->nty, c->ntd, c->nslots}; }

// True when a cell lies outside [1, N-3] on either axis (outflow).
__device__ __forceinline__ bool cell_out(const EGeo& g, int ci, int cj) {
    return ci < 1 || ci > g.nx - 3 || cj < 1 || cj > g.ny - 3;
}

// Sort key of a particle: its tile -- live (slot, tile), the parking tile of
// its slot (the slot finished, or the particle flowed out), or -1 (released
// slot: the particle drops out of the next sort) -- and, in *cb, its cell
// in the tile (the stencil base, (ib & 15) << 4 | (jb & 15)).  Tile ids use the
// whole int32 range, so the number of slots on a terrain is bounded by
// slots x terrain tiles < 2^31.
__device__ __forceinline__ int ekey(const EGeo& g, int st, int slot, bool out, int cell, float ox, float oy,
                                    int* tile, int* cb) {
    *cb = 0;
    if (slot_active(st) && !out) {
        const int ci = cell >> 16, cj = cell & 0xffff;
        const int ib = ebase(ci, ox, g.nx - 3), jb = ebase(cj, oy, g.ny - 3);
        const int t = slot * g.ntd + (ib >> 4) * g.nty + (jb >> 4);
        *tile = t;
        *cb = ((ib & 15) << 4) | (jb & 15);
        return t;
    }
    if (slot_active(st) || st == SLOT_DONE || st == SLOT_FAILED) {
        const int t = g.nslots * g.ntd + slot;
        *tile = t;
        return t;
    }
    *tile = 0;
    return -1;
}

// Per-slot max |v|^2 as a sharded integer max of the IEEE bits (|v|^2 >= 0;
// NaN sorts above +inf).  Particles are stored by slot, so a wave almost
// always holds one slot: it reduces across lanes and issues one atomic on its
// shard (wave index % VMAX_SHARDS); a wave that straddles slots issues one
// atomic per lane.  Every lane of the wave must call this.
// The wave maximum is taken over the FP32 bits (|v|^2 >= 0 orders like its
// bits, NaN above +inf) and widened to the FP64 bits the shards hold, which
// is the maximum of the widened values.
__device__ __forceinline__ void publish_vmax(int slot, bool valid, float v2, unsigned long long* __restrict__ vmx) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned bits32 = valid ? __float_as_uint(v2) : 0u;
    const unsigned long long vmask = __ballot(valid);
    const int first = vmask ? __ffsll((long long)vmask) - 1 : 0;
    const int s0 = __shfl(slot, first, 64);
    const unsigned long long same = __ballot(valid && slot == s0);
    const unsigned m32 = wave_max_u32(bits32);
    const unsigned sh = (blockIdx.x * (WG / 64) + wv) & (VMAX_SHARDS - 1);
    if (same == vmask) {
        if (lane == 0 && vmask)
            atomicMax(&vmx[(size_t)s0 * VMAX_SHARDS + sh],
                      (unsigned long long)__double_as_longlong((double)__uint_as_float(m32)));
    } else if (valid) {
        atomicMax(&vmx[(size_t)slot * VMAX_SHARDS + sh], (unsigned long long)__double_as_longlong((double)v2));
    }
}

// Counts each lane's particle into its tile: one atomic per run of equal tiles
// among consecutive valid lanes, and no value returned (the slot inside the
// tile is taken later, by the scatter).  Every lane of the wave must call this.
__device__ __forceinline__ void wave_count_any(int t, bool valid, int* __restrict__ tile_count) {
    const int lane = threadIdx.x & 63;
    const int tprev = lane_prev(t);
    const bool vprev = lane_prev((int)valid) != 0;
    const bool head = valid && (lane == 0 || !vprev || tprev != t);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long hmask = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long brk = (hmask | ~vmask) & ~upto; // next head or next invalid lane
    const int end = brk ? (__ffsll((long long)brk) - 1) : 64;
    if (head) atomicAdd(&tile_count[t], end - lane);
}

// --- control ---------------------------------------------------------------

__global__ void k_eclear(ECtl* c, Slot* __restrict__ slots) {
    if (threadIdx.x == 0) {
        if (c->stop == STOP_FRAME || c->stop == STOP_DONE) c->stop = RUNNING;
        c->steps_in_call = 0;
    }
    for (int s = threadIdx.x; s < c->nslots; s += blockDim.x) {
        slots[s].frame_hit = 0;
        slots[s].steps_in_call = 0;
    }
}

// Start of a step: per slot, the while condition (src/simulation.cpp:500) and
// updateDt (565-582) from the slot's own max |v|; priming slots get dt = 0
// (simulation.hpp:39, simulation.cpp:493-495).
// Each slot's dt shards -> shard 0, the others cleared: one wave per slot.
static_assert(VMAX_SHARDS % 64 == 0, "shards per lane");
// The maximum of a slot's shards, in every lane of the (whole) wave.
__device__ __forceinline__ unsigned long long vmax_shards(const unsigned long long* __restrict__ b) {
    const int lane = threadIdx.x & 63;
    unsigned long long q[VMAX_SHARDS / 64];
#pragma unroll
    for (int j = 0; j < VMAX_SHARDS / 64; ++j) q[j] = b[lane + 64 * j]; // all in flight together
    unsigned long long v = 0;
#pragma unroll
    for (int j = 0; j < VMAX_SHARDS / 64; ++j) v = q[j] > v ? q[j] : v;
    return wave_max_u64(v);
}
__device__ __forceinline__ void vmax_slot(unsigned long long* __restrict__ b) {
    const int lane = threadIdx.x & 63;
    const unsigned long long v = vmax_shards(b);
#pragma unroll
    for (int j = 0; j < VMAX_SHARDS / 64; ++j) b[lane + 64 * j] = (lane == 0 && j == 0) ? v : 0ull;
}

// Batches: the shard reduction of every slot, one wave (block) per slot.
__global__ void __launch_bounds__(64) k_evmax(const ECtl* c, unsigned long long* __restrict__ vmx) {
    if (c->stop || c->compact) return;
    vmax_slot(vmx + (size_t)blockIdx.x * VMAX_SHARDS);
}

// End of a started step and start of the next one, run by one workgroup: the
// extra block of k_ereduce (the tile counting does not read what this writes:
// its counts come from the previous G2P), or k_eend after the last step the
// host enqueued (begin = false).
//  * End (pend_end), src/simulation.cpp:503-509: time += dt and the frame
//    trigger per running slot; priming slots start running.  A single run
//    stops at a frame (the caller writes its outputs), so no step starts.
//  * Start (begin): per slot, the while condition (src/simulation.cpp:500)
//    and updateDt (565-582) from the slot's own max |v|; priming slots get
//    dt = 0 (simulation.hpp:39, simulation.cpp:493-495).
//  * A re-sort pass (compact) ends by flipping the buffers and starts with no
//    slot advancing.
// Every control and slot field is read before the block's first store, so
// the reads are issued together (the control block through the scalar cache)
// instead of one dependent reload after each store.
struct ECtlIn {
    int pend, compact, nslots, single, reorder, parity, n_sorted, n_active, sic, epoch, every;
    double dx;
};
__device__ __forceinline__ ECtlIn ectl_in(const ECtl* c) {
    return {c->pend_end, c->compact, c->nslots, c->single, c->reorder, c->parity,
            c->n_sorted, c->n_active, c->steps_in_call, c->epoch, c->reorder_every, c->dx};
}

// Sum over the workgroup, in every thread (every thread calls it; a caller
// that calls it again first passes a barrier, as its LDS is reused).
__device__ __forceinline__ int block_sum_i32(int v) {
    __shared__ int s_w[WG / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < WG / 64; ++w) t += s_w[w];
    return t;
}

// A slot's fields the end and start of a step read.
struct ESlotIn {
    int status, frame, sic, step, keep;
    double time, dt, output_dt, total, cfl, dt_max;
};
__device__ __forceinline__ ESlotIn eslot_in(const Slot& S) {
    return {S.status, S.frame, S.steps_in_call, S.step, S.keep_grid,
            S.time, S.dt, S.output_dt, S.total_time, S.C_cfl, S.dt_max};
}

// End of a started step and start of the next one, run by one workgroup: the
// extra block of k_ereduce (the tile counting does not read what this writes:
// its counts come from the previous G2P), or k_eend after the last step the
// host enqueued (begin = false).
//  * End (pend_end), src/simulation.cpp:503-509: time += dt and the frame
//    trigger per running slot; priming slots start running.  A single run
//    stops at a frame (the caller writes its outputs), so no step starts.
//  * Start (begin): per slot, the while condition (src/simulation.cpp:500)
//    and updateDt (565-582) from the slot's own max |v|; priming slots get
//    dt = 0 (simulation.hpp:39, simulation.cpp:493-495).
//  * A re-sort pass (compact) ends by flipping the buffers and starts with no
//    slot advancing.
// The control fields, the slot fields and the max |v| shards are all read
// before the block's first store and before any branch on them, so they are
// in flight together instead of a dependent reload after each store.
// Slot ownership: with k_evmax (reduced), thread s owns slots s, s + WG, ...;
// otherwise (at most 4 slots) lane 0 of wave s owns slot s, whose shards its
// wave reduces.  nsl is the slot count (from the terrain geometry).
__device__ void ectl_block(ECtl* c, int nsl, Slot* __restrict__ slots, unsigned long long* __restrict__ vmx,
                           int reduced, bool begin) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int s_first = reduced ? (int)threadIdx.x : (lane == 0 && wv < nsl ? wv : nsl);
    const int s_step = reduced ? WG : nsl;
    // first owned slot's fields and shards, read unconditionally (clamped to
    // a valid slot) so that they are in flight with the control fields
    const int s_pre = min(s_first, nsl - 1);
    ESlotIn q0 = eslot_in(slots[s_pre]);
    const unsigned long long vr0 = (begin && reduced) ? vmx[(size_t)s_pre * VMAX_SHARDS] : 0ull;
    unsigned long long vsh = 0; // this wave's slot max (used when not reduced)
    if (begin) vsh = vmax_shards(vmx + (size_t)min(wv, nsl - 1) * VMAX_SHARDS);
    const ECtlIn in = ectl_in(c);
    const int pend = in.pend, compact = in.compact, single = in.single;
    // the step that starts is a re-sort pass (a pending pass ends first)
    const bool start_compact = !pend && compact;
    const bool slot_steps = begin && !start_compact;
    const bool slot_ends = pend && !compact;
    const int ep = in.epoch + 1;
    int hits = 0, act = 0, outs = 0;
    for (int s = s_first; s < nsl; s += s_step) {
        Slot& S = slots[s];
        ESlotIn q = q0;
        unsigned long long vr = vr0;
        if (s != s_first) {
            q = eslot_in(S);
            vr = (begin && reduced) ? vmx[(size_t)s * VMAX_SHARDS] : 0ull;
        }
        if (!slot_ends && !slot_steps) continue;
        bool hit = false;
        if (slot_ends) {
            if (q.status == SLOT_PRIMING) {
                q.status = SLOT_RUNNING;
                S.status = q.status;
            } else if (q.status == SLOT_RUNNING) {
                q.time += q.dt;
                S.time = q.time;
                S.steps_in_call = q.sic + 1;
                if (q.output_dt * (double)(q.frame + 1) - q.time < 1e-13) {
                    q.frame += 1;
                    S.frame = q.frame;
                    S.frame_hit = 1;
                    hit = true;
                    ++hits;
                }
            }
        }
        // a single run holds one slot: its frame stops the run before a step starts
        if (!slot_steps || (single && hit)) continue;
        if (reduced) vmx[(size_t)s * VMAX_SHARDS] = 0ull;
        if (q.status == SLOT_PRIMING) {
            S.dt = 0.0;
            S.out_step = 1;
            S.vz_due = 1;
            S.last_out_epoch = ep;
            ++act;
            continue;
        }
        if (q.status != SLOT_RUNNING) continue;
        if (!(q.time < q.total)) {
            S.status = SLOT_DONE;
            continue;
        }
        const double v2 = __longlong_as_double(reduced ? vr : vsh);
        if (!(v2 <= 1.7976931348623157e308)) {
            S.status = SLOT_FAILED;
            S.err = 5; // SLAB_ENAN
            continue;
        }
        const double max_speed = sqrt(v2);
        double dt;
        if (max_speed > 0.0) {
            const double a = q.cfl * in.dx / max_speed;
            dt = (q.dt_max < a) ? q.dt_max : a;
        } else {
            dt = q.dt_max;
        }
        const double next_frame_time = q.output_dt * (double)(q.frame + 1);
        if (q.time + dt > next_frame_time - 1e-12) dt = next_frame_time - q.time;
        S.dt = dt;
        S.step = q.step + 1;
        const double tnew = q.time + dt;
        const int hit2 = (q.output_dt * (double)(q.frame + 1) - tnew < 1e-13);
        const int out = hit2 || q.keep;
        S.out_step = out;
        // particle vz feeds only output steps (P2G's grid vz) and the particle
        // output of a frame: G2P refreshes it on output steps and on steps
        // after which the next step can reach its frame (its dt is at most
        // dt_max, and is clipped to land on the frame within 1e-12)
        const double next_frame = q.output_dt * (double)(q.frame + hit2 + 1);
        S.vz_due = out || (tnew + q.dt_max >= next_frame - 1e-12);
        if (out) {
            S.last_out_epoch = ep;
            outs = 1;
        }
        ++act;
    }
    const int n_hit = block_sum_i32(hits);
    __syncthreads(); // block_sum_i32's LDS is reused below
    const int n_act = block_sum_i32(act);
    __syncthreads();
    const int n_out = block_sum_i32(outs);
    const bool stop_frame = slot_ends && single && n_hit > 0;
    // the shards of the slots (at most 4) that start a step are cleared
    if (slot_steps && !stop_frame && !reduced && wv < nsl) {
        unsigned long long* sh = vmx + (size_t)wv * VMAX_SHARDS;
#pragma unroll
        for (int j = 0; j < VMAX_SHARDS / 64; ++j) sh[lane + 64 * j] = 0ull;
    }
    if (threadIdx.x == 0) {
        int par = in.parity;
        if (pend) {
            if (compact) {
                par ^= 1;
                c->n_store = in.n_sorted;
                c->compact = 0;
                c->reorder = 0;
            } else {
                if (in.reorder) {
                    par ^= 1;
                    c->n_store = in.n_sorted;
                }
                if (in.n_active > 0) c->steps_in_call = in.sic + 1;
                if (stop_frame) c->stop = STOP_FRAME;
            }
            c->parity = par;
            c->pend_end = 0;
        }
        if (begin && !stop_frame) {
            c->epoch = ep;
            c->p2g_parity = par;
            if (start_compact) {
                c->reorder = 1;
                c->pend_end = 1;
            } else {
                c->n_active = n_act;
                // a single run keeps the pre-G2P positions of output steps (its
                // remesh window is read from them), so those steps write the
                // other buffer
                c->reorder = in.every <= 1 || ep % in.every == 0 || (single && n_out > 0);
                if (n_act == 0)
                    c->stop = STOP_DONE; // nothing to advance
                else
                    c->pend_end = 1; // ended by the next step's first kernel or by k_eend
            }
        }
    }
}

__global__ void __launch_bounds__(WG) k_eend(ECtl* c, Slot* __restrict__ slots) {
    if (c->stop || !c->pend_end) return;
    ectl_block(c, c->nslots, slots, nullptr, 1, false);
}

// --- keys of admitted particles ---------------------------------------------

__global__ void __launch_bounds__(WG) k_ekeys(ECtl* c, const Slot* __restrict__ slots, PB32 A, PB32 B, int s0,
                                              int s1, int* __restrict__ key, uint8_t* __restrict__ kcell,
                                              int* __restrict__ tile_count) {
    const PB32 P = ecur(c, A, B);
    const int s = s0 + blockIdx.x * WG + threadIdx.x;
    const bool in = s < s1;
    int t = 0, k = -1, cb = 0;
    if (in) {
        const int raw = P.slot[s], sl = raw & SLOT_MASK;
        const float2 o = P.o[s];
        k = ekey(egeo(c), slots[sl].status, sl, (raw & SLOT_OUT) != 0, P.cell[s], o.x, o.y, &t, &cb);
    }
    wave_count_any(t, in && k >= 0, tile_count);
    if (in) {
        key[s] = k;
        kcell[s] = (uint8_t)cb;
    }
}

// Initial particle z, src/simulation.cpp:455-465 (4x4 quadratic weights).
__global__ void k_ez(ECtl* c, PB32 A, PB32 B, int s0, int s1, const float* __restrict__ tz) {
    const PB32 P = ecur(c, A, B);
    const int s = s0 + blockIdx.x * WG + threadIdx.x;
    if (s >= s1) return;
    const int cell = P.cell[s];
    const int ci = cell >> 16, cj = cell & 0xffff;
    const float2 o = P.o[s];
    const float ox = o.x, oy = o.y;
    float z = 0.0f;
    for (int i = ci - 1; i < ci + 3; ++i) {
        for (int j = cj - 1; j < cj + 3; ++j) {
            if (i < 0 || j < 0 || i >= c->nx || j >= c->ny) continue;
            z += tz[(size_t)i * c->ny + j] * (Nf((float)(ci - i) + ox) * Nf((float)(cj - j) + oy));
        }
    }
    P.z[s] = z;
}

// --- tile scan over (slot, tile) + parking tiles ------------------------------

// Flags of a thread's SCAN_ITEMS consecutive tiles t0 + k: particles, is a
// particle tile, is a grid tile (it or its left / lower / diagonal neighbour
// holds particles).  Every count is read first, at clamped in-range indices
// (unused where the tile has no such neighbour), so the reads are in flight
// together.
__device__ __forceinline__ void etile_flags(const EGeo& geo, const int* __restrict__ tile_count, int t0, int ntiles,
                                            int f[SCAN_ITEMS][3]) {
    const int nty = geo.nty, live = geo.nslots * geo.ntd;
    int n[SCAN_ITEMS], nl[SCAN_ITEMS], nb[SCAN_ITEMS], nd[SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int t = min(t0 + k, ntiles - 1);
        n[k] = tile_count[t];
        nl[k] = tile_count[max(t - nty, 0)];
        nb[k] = tile_count[max(t - 1, 0)];
        nd[k] = tile_count[max(t - nty - 1, 0)];
    }
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int t = t0 + k;
        f[k][0] = f[k][1] = f[k][2] = 0;
        if (t >= ntiles) continue;
        f[k][0] = n[k];
        f[k][1] = n[k] > 0;
        if (t >= live) continue; // parking tile: particles, never a grid tile
        const int local = t % geo.ntd;
        const int tx = local / nty, ty = local - tx * nty;
        int g = n[k] > 0;
        if (tx > 0) g |= nl[k] > 0;
        if (ty > 0) g |= nb[k] > 0;
        if (tx > 0 && ty > 0) g |= nd[k] > 0;
        f[k][2] = g;
    }
}

__device__ __forceinline__ void block_exscan3e(int v[3], int out[3], int tot[3]) {
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

// True when the block's tiles all belong to one FREE slot: a released slot's
// tiles hold no particles (release re-keys and compacts), so the block has
// nothing to count or list.
__device__ __forceinline__ bool scan_block_idle(const EGeo& geo, const Slot* __restrict__ slots, int ntiles) {
    const int t0 = blockIdx.x * SCAN_BLOCK, t1 = min(t0 + SCAN_BLOCK, ntiles) - 1;
    const int s0 = t0 / geo.ntd, s1 = t1 / geo.ntd;
    return s0 == s1 && s0 < geo.nslots && slots[s0].status == SLOT_FREE;
}

// Per scan block: tile counts -> the block's numbers of particles, particle
// tiles and grid tiles.  The last block ends the previous step if it is
// pending and starts this one (ectl_block); the counting reads
// only the tile counts and which slots are FREE, neither of which the end or
// start of a step changes.
__global__ void __launch_bounds__(WG) k_ereduce(ECtl* c, Slot* __restrict__ slots,
                                                const int* __restrict__ tile_count, int ntiles,
                                                int* __restrict__ bsum, unsigned long long* __restrict__ vmx,
                                                int reduced) {
    if (c->stop) return;
    // control fields read before any store on any path
    const EGeo geo = egeo(c);
    if (blockIdx.x == gridDim.x - 1) {
        ectl_block(c, geo.nslots, slots, vmx, reduced, true);
        return;
    }
    if (scan_block_idle(geo, slots, ntiles)) {
        if (threadIdx.x < 3) bsum[3 * blockIdx.x + threadIdx.x] = 0;
        return;
    }
    int acc[3] = {0, 0, 0};
    const int t0 = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    int f[SCAN_ITEMS][3];
    etile_flags(geo, tile_count, t0, ntiles, f);
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        acc[0] += f[k][0];
        acc[1] += f[k][1];
        acc[2] += f[k][2];
    }
    int ex[3], tot[3];
    block_exscan3e(acc, ex, tot);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) bsum[3 * blockIdx.x + k] = tot[k];
}

// The totals of k_escan (particles, particle tiles, grid tiles) into the
// control block; an active-tile count above the capacity (tile_cap, gtile_cap:
// read by the caller before its first store) stops the step with an error.
// Thread 0 of one block.
__device__ __forceinline__ void scan_totals(ECtl* c, const int tot[3], int tile_cap, int gtile_cap) {
    c->n_sorted = tot[0];
    c->n_ptiles = tot[1];
    c->n_gtiles = tot[2];
    if (tot[1] > tile_cap || tot[2] > gtile_cap) {
        c->err = 7;
        c->stop = STOP_ERROR;
    }
}

__global__ void __launch_bounds__(WG) k_escan(ECtl* c, const int* __restrict__ bsum, int nblk,
                                              int* __restrict__ boff) {
    if (c->stop) return;
    const int tcap = c->tile_cap, gcap = c->gtile_cap;
    int run[3] = {0, 0, 0};
    for (int b0 = 0; b0 < nblk; b0 += WG) {
        const int b = b0 + threadIdx.x;
        int v[3] = {0, 0, 0};
        if (b < nblk)
            for (int k = 0; k < 3; ++k) v[k] = bsum[3 * b + k];
        int ex[3], tot[3];
        block_exscan3e(v, ex, tot);
        if (b < nblk)
            for (int k = 0; k < 3; ++k) boff[3 * b + k] = run[k] + ex[k];
        for (int k = 0; k < 3; ++k) run[k] += tot[k];
    }
    if (threadIdx.x == 0) scan_totals(c, run, tcap, gcap);
}

// Per scan block: tile begin offsets and the particle / grid tile lists.  The
// block's starting offsets come from k_escan (boff), or, with few scan blocks
// (merged != 0), from block_prefix3, block 0 then writing the totals.
__global__ void __launch_bounds__(WG) k_eoffsets(ECtl* c, const Slot* __restrict__ slots,
                                                 const int* __restrict__ tile_count, int ntiles,
                                                 const int* __restrict__ boff, const int* __restrict__ bsum,
                                                 int nblk, int merged, int* __restrict__ tile_begin,
                                                 int* __restrict__ ptiles, int* __restrict__ gtiles,
                                                 int* __restrict__ pstamp, int* __restrict__ gstamp,
                                                 int* __restrict__ tile_pos, int* __restrict__ gtile_pos,
                                                 int* __restrict__ hm_map, const int* __restrict__ hm_stack,
                                                 int2* __restrict__ pinfo) {
    if (c->stop) return;
    // control-block fields read before this kernel's first store
    const EGeo geo = egeo(c);
    const int ep = c->epoch, tcap = c->tile_cap, gcap = c->gtile_cap;
    const bool idle = scan_block_idle(geo, slots, ntiles);
    if (idle && blockIdx.x != 0) return; // block 0 writes the totals even when idle
    int base[3];
    if (merged) {
        int tot[3];
        block_prefix3(bsum, nblk, base, tot);
        if (blockIdx.x == 0 && threadIdx.x == 0) scan_totals(c, tot, tcap, gcap);
        if (tot[1] > tcap || tot[2] > gcap) return; // every block sees the same totals
    } else {
        for (int q = 0; q < 3; ++q) base[q] = boff[3 * blockIdx.x + q];
    }
    if (idle) return;
    const int t0 = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_ITEMS;
    int f[SCAN_ITEMS][3];
    int acc[3] = {0, 0, 0};
    etile_flags(geo, tile_count, t0, ntiles, f);
    for (int k = 0; k < SCAN_ITEMS; ++k)
        for (int q = 0; q < 3; ++q) acc[q] += f[k][q];
    int ex[3], tot[3];
    block_exscan3e(acc, ex, tot);
    int run[3];
    for (int q = 0; q < 3; ++q) run[q] = base[q] + ex[q];
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int t = t0 + k;
        if (t >= ntiles) break;
        tile_begin[t] = run[0];
        if (f[k][1]) {
            ptiles[run[1]] = t;
            pinfo[run[1]] = make_int2(run[0], f[k][0]); // the tile's begin and count, by list position
            tile_pos[t] = run[1];
            pstamp[t] = ep;
        }
        if (f[k][2]) {
            gtiles[run[2]] = t;
            gtile_pos[t] = run[2];
            gstamp[t] = ep;
            // a batch's max maps live in a pool of the tiles its slots have
            // touched: a grid tile takes a (zeroed) pool tile the first time
            if (hm_map && hm_map[t] < 0) {
                const int q = atomicAdd(&c->hm_top, 1);
                if (q < c->hm_cap) {
                    hm_map[t] = hm_stack[q];
                } else {
                    c->err = 8;
                    c->stop = STOP_ERROR;
                }
            }
        }
        for (int q = 0; q < 3; ++q) run[q] += f[k][q];
    }
}

// --- sort ------------------------------------------------------------------

// Buckets the stored particles by tile: each takes the next slot of its tile
// from a per-tile cursor (the sort clears the cursors of the tiles it orders).
// A wave handles SC_PER groups of 64 consecutive particles and issues the
// cursor atomics of all groups (one per run of equal tiles) before it waits
// for any of them.  Arrival order inside a tile is arbitrary: the sort makes
// it canonical.
constexpr int SC_PER = 16;
constexpr int SC_HASH = 512; // per-workgroup table of the tiles its particles go to
__global__ void __launch_bounds__(WG) k_escatter(ECtl* c, const int* __restrict__ key,
                                                 const int* __restrict__ tile_begin, int* __restrict__ cursor,
                                                 int* __restrict__ tmp) {
    if (c->stop) return;
    // The workgroup's 4,096 consecutive particles take their slots per tile in
    // two levels: each run of equal tiles in a 64-particle group adds its
    // length to its tile's entry of an LDS table (offset within the
    // workgroup), then each table entry takes the workgroup's range of its
    // tile with one cursor atomic.  Storage is canonical at its last re-sort,
    // so a workgroup's particles go to few tiles: a few global atomics per
    // workgroup instead of one per run.  A run whose tile finds no table entry
    // (more than SC_HASH tiles in the workgroup) takes its range directly.
    __shared__ int s_ht[SC_HASH]; // tile id, or -1
    __shared__ int s_hc[SC_HASH]; // particles, then the workgroup's base in the tile
    for (int i = threadIdx.x; i < SC_HASH; i += WG) {
        s_ht[i] = -1;
        s_hc[i] = 0;
    }
    const int n = c->n_store, lane = threadIdx.x & 63;
    const int s0 = (blockIdx.x * WG + (threadIdx.x & ~63)) * SC_PER + lane;
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    int t[SC_PER], start[SC_PER], hv[SC_PER];
    bool valid[SC_PER];
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) {
        const int s = s0 + 64 * j;
        const int k = s < n ? key[s] : -1;
        valid[j] = k >= 0;
        t[j] = valid[j] ? k : 0;
    }
    // the tiles' begins and the previous lane's tile of every group first
    // (all in flight together), then the run heads
    int tb[SC_PER], tprev[SC_PER];
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) tb[j] = valid[j] ? tile_begin[t[j]] : 0;
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) tprev[j] = __shfl_up(valid[j] ? t[j] : -1, 1, 64);
    __syncthreads(); // the table is cleared
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) {
        const bool head = valid[j] && (lane == 0 || tprev[j] != t[j]);
        const unsigned long long vmask = __ballot(valid[j]);
        const unsigned long long hmask = __ballot(head);
        const unsigned long long hle = hmask & upto;
        const unsigned long long brk = (hmask | ~vmask) & ~upto;
        start[j] = hle ? 63 - __clzll((long long)hle) : 0;
        const int end = brk ? (__ffsll((long long)brk) - 1) : 64;
        hv[j] = 0;
        if (head) {
            // hv >= 0: table entry << 16 | offset in the workgroup's range;
            // hv < 0: -1 - the run's base in the tile (taken directly)
            unsigned h = ((unsigned)t[j] * 2654435761u) >> 23; // top 9 bits
            int e = -1;
            for (int probe = 0; probe < SC_HASH; ++probe) {
                const int old = atomicCAS(&s_ht[h], -1, t[j]);
                if (old == -1 || old == t[j]) {
                    e = (int)h;
                    break;
                }
                h = (h + 1) & (SC_HASH - 1);
            }
            hv[j] = e >= 0 ? (e << 16) | atomicAdd(&s_hc[e], end - lane) : -1 - atomicAdd(&cursor[t[j]], end - lane);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SC_HASH; i += WG)
        if (s_ht[i] >= 0) s_hc[i] = atomicAdd(&cursor[s_ht[i]], s_hc[i]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SC_PER; ++j) {
        const int hj = __shfl(hv[j], start[j], 64);
        const int b = hj >= 0 ? s_hc[hj >> 16] + (hj & 0xffff) : -1 - hj;
        if (valid[j]) tmp[tb[j] + b + (lane - start[j])] = s0 + 64 * j;
    }
}

// Per particle tile: counting sort by cell, then each cell's run sorted by
// creation index: canonical order (slot, tile, cell, creation index).
//  * Tiles of at most CAP particles stage (cell, creation index) in LDS with
//    one coalesced pass.  If no cell holds more than ES_RANK particles, each
//    particle's place in its cell is its rank there: the number of the cell's
//    particles with a smaller creation index (unique within a slot), counted
//    by the particle's own thread, so no cell is sorted serially; otherwise
//    the whole tile is bitonic-sorted on 64-bit keys (cell, creation index,
//    arrival position).
//  * Larger tiles place their particles by cell in global memory, then sort
//    the cells in rounds: each round bitonic-sorts a contiguous range of whole
//    cells (at most CAP / 2 particles) in LDS; only a single cell of more than
//    CAP / 2 particles sorts in global memory.
//  * Parking tiles keep their bucket order (their particles are only carried).
// CAP, the in-LDS sort capacity, is a launch-shape choice (1,024, 2,048 or
// 4,096 particles): a smaller one holds fewer registers and LDS per workgroup
// (more workgroups per CU) but sends more tiles through the rounds; the
// autotuner picks it on the live workload.  The order does not depend on it.
constexpr int ES_RANK = 256;
constexpr int ES_GROUP = 4; // particles per thread whose global reads are issued together
constexpr int ES_CAPS[3] = {1024, 2048, 4096};

// Ascending bitonic sort of k[0, P), P a power of two, by one workgroup.
__device__ __forceinline__ void bitonic_lds(unsigned long long* k, int P) {
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < (P >> 1); i += WG) {
                const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const unsigned long long x = k[lo], y = k[hi];
                if ((x > y) == asc) {
                    k[lo] = y;
                    k[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ int pow2_ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// The thread index through an empty asm: values derived from it are computed
// where they are used instead of being hoisted out of a loop and held in
// registers across all of it.
__device__ __forceinline__ int tid_opaque() {
    int v = threadIdx.x;
    asm volatile("" : "+v"(v));
    return v;
}

// LDS of the tile sort: histogram, cell offsets, scan and round scratch, and
// the staging region (declared once by the kernel, shared by its passes).
struct ESortLds {
    int* hist;
    int* off;
    int* w4;
    int* c1;
    int* cmax;
    unsigned char* mem;
};

template <int CAP>
__device__ __forceinline__ void esort_tiles(ECtl* c, PB32 A, PB32 B, const uint8_t* __restrict__ kcell,
                                            const int* __restrict__ ptiles, const int2* __restrict__ pinfo,
                                            int* __restrict__ tile_count, int* __restrict__ cursor,
                                            const int* __restrict__ tmp, int* __restrict__ idx,
                                            int* __restrict__ cell_begin, const uint8_t* __restrict__ only,
                                            const ESortLds& L) {
    constexpr int ROUND = CAP / 2; // keys and storage indices of a round fit the staging area
    static_assert(CAP % (WG * ES_GROUP) == 0 && CAP <= 8192, "sort capacity");
    int* s_hist = L.hist;
    int* s_off = L.off;
    int* s_w = L.w4;
    int& s_c1 = *L.c1;
    int& s_cmax = *L.cmax;
    unsigned char* s_mem = L.mem;
    // one LDS region (s_mem, CAP * 9 bytes), three uses: (a) creation index by
    // arrival, creation index in cell order and cell of a small tile; (b)
    // 64-bit keys of a small tile; (c) keys + storage index of a round
    int* s_pid = reinterpret_cast<int*>(s_mem);                                    // (a) CAP x i32
    int* s_pidc = reinterpret_cast<int*>(s_mem + CAP * 4);                         // (a) CAP x i32
    unsigned char* s_cell = s_mem + CAP * 8;                                       // (a) CAP x u8
    unsigned long long* s_key = reinterpret_cast<unsigned long long*>(s_mem);      // (b) CAP, (c) ROUND
    int* s_sv = reinterpret_cast<int*>(s_mem + ROUND * 8);                         // (c) ROUND x i32
    const int32_t* __restrict__ pid = ecur(c, A, B).pid;
    const int nt = c->n_ptiles, live = c->nslots * c->ntd;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        if (!only[w]) continue; // the tiles the small-tile pass left
        // the tile, its begin and its count, by list position: read-only here,
        // so read together through the scalar cache
        const int t = ptiles[w];
        const int2 nb = pinfo[w];
        const int n = nb.y, b = nb.x; // n >= 1: listed tiles hold particles
        // the first group of bucket reads and their key and creation-index
        // reads touch no LDS: issued before the barrier, they overlap it
        int sv0[ES_GROUP], kc0[ES_GROUP], pd0[ES_GROUP];
#pragma unroll
        for (int j = 0; j < ES_GROUP; ++j) sv0[j] = tmp[b + min((int)threadIdx.x + j * WG, n - 1)];
#pragma unroll
        for (int j = 0; j < ES_GROUP; ++j) {
            kc0[j] = kcell[sv0[j]];
            pd0[j] = pid[sv0[j]];
        }
        if (threadIdx.x == 0) s_cmax = 0;
        s_hist[threadIdx.x] = 0;
        __syncthreads();
        int* cb = cell_begin + (size_t)w * CELL_STRIDE;
        if (t >= live) { // parking tile
            for (int r = threadIdx.x; r < n; r += WG) idx[b + r] = tmp[b + r];
            cb[threadIdx.x] = threadIdx.x == 0 ? b : b + n;
            if (threadIdx.x == 0) {
                cb[TILE_CELLS] = b + n;
                tile_count[t] = 0;
                cursor[t] = 0;
            }
            __syncthreads();
            continue;
        }
        const bool small = n <= CAP;
        // storage index of each arrival this thread stages (small tiles): the
        // final write places it from registers instead of re-reading tmp
        int svr[CAP / WG];
        if (small) {
            // groups of ES_GROUP arrivals per thread: the bucket reads of a group
            // are issued together, then their key and creation-index reads
            // (clamped indices keep every read in the tile); the next group's
            // bucket reads are issued before this group's keys are waited for
            int svn[ES_GROUP];
#pragma unroll
            for (int j = 0; j < ES_GROUP; ++j)
                svn[j] = ES_GROUP * WG < n ? tmp[b + min((int)threadIdx.x + (ES_GROUP + j) * WG, n - 1)] : 0;
#pragma unroll
            for (int k0 = 0; k0 < CAP / WG; k0 += ES_GROUP) {
                if (k0 * WG < n) {
                    int sv[ES_GROUP], kc[ES_GROUP], pd[ES_GROUP];
                    if (k0 == 0) {
#pragma unroll
                        for (int j = 0; j < ES_GROUP; ++j) {
                            sv[j] = sv0[j];
                            kc[j] = kc0[j];
                            pd[j] = pd0[j];
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < ES_GROUP; ++j) sv[j] = svn[j];
                        if (k0 + ES_GROUP < CAP / WG && (k0 + ES_GROUP) * WG < n) {
#pragma unroll
                            for (int j = 0; j < ES_GROUP; ++j)
                                svn[j] = tmp[b + min((int)threadIdx.x + (k0 + ES_GROUP + j) * WG, n - 1)];
                        }
#pragma unroll
                        for (int j = 0; j < ES_GROUP; ++j) {
                            kc[j] = kcell[sv[j]];
                            pd[j] = pid[sv[j]];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) {
                        const int r = threadIdx.x + (k0 + j) * WG;
                        svr[k0 + j] = sv[j];
                        if (r < n) {
                            const int cc = kc[j];
                            s_cell[r] = (unsigned char)cc;
                            s_pid[r] = pd[j];
                            atomicAdd(&s_hist[cc], 1);
                        }
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) svr[k0 + j] = 0;
                }
            }
        } else {
            for (int r0 = threadIdx.x; r0 < n; r0 += WG * ES_GROUP) {
                int kc[ES_GROUP];
                if (r0 == (int)threadIdx.x) { // the first group was read before the barrier
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) kc[j] = kc0[j];
                } else {
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) kc[j] = kcell[tmp[b + min(r0 + j * WG, n - 1)]];
                }
#pragma unroll
                for (int j = 0; j < ES_GROUP; ++j)
                    if (r0 + j * WG < n) atomicAdd(&s_hist[kc[j]], 1);
            }
        }
        __syncthreads();
        const int cnt = s_hist[threadIdx.x];
        if (cnt > ES_RANK) atomicMax(&s_cmax, cnt);
        {
            const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
            int inc = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                int u = __shfl_up(inc, o, 64);
                if (lane >= o) inc += u;
            }
            if (lane == 63) s_w[wv] = inc;
            __syncthreads();
            int pre = 0;
            for (int q = 0; q < wv; ++q) pre += s_w[q];
            s_off[threadIdx.x] = pre + inc - cnt;
            cb[threadIdx.x] = b + pre + inc - cnt;
            if (threadIdx.x == 0) {
                cb[TILE_CELLS] = b + n;
                tile_count[t] = 0;
                cursor[t] = 0;
            }
        }
        __syncthreads();
        const int off = s_off[threadIdx.x];
        s_hist[threadIdx.x] = off; // cursor
        const bool dense = s_cmax > ES_RANK;
        __syncthreads();
        if (small && !dense) {
            for (int r = threadIdx.x; r < n; r += WG) s_pidc[atomicAdd(&s_hist[s_cell[r]], 1)] = s_pid[r];
            __syncthreads();
            // s_pidc[s_off[c] .. s_hist[c]) now holds cell c's creation indices;
            // each arrival is placed from its registers at the cell start plus
            // its rank (creation indices are unique within the tile's slot)
#pragma unroll
            for (int k = 0; k < CAP / WG; ++k) {
                const int r = threadIdx.x + k * WG;
                if (r < n) {
                    const int cr = s_cell[r], pr = s_pid[r];
                    const int lo = s_off[cr], hi = s_hist[cr];
                    int rank = 0;
                    for (int q = lo; q < hi; ++q) rank += s_pidc[q] < pr ? 1 : 0;
                    idx[b + lo + rank] = svr[k];
                }
            }
        } else if (small) {
            // keys (cell << 44 | creation index << 13 | arrival position), built
            // in registers first because they overwrite the staged arrays
            unsigned long long kv[CAP / WG];
#pragma unroll
            for (int k = 0; k < CAP / WG; ++k) {
                const int r = threadIdx.x + k * WG;
                kv[k] = r < n ? ((unsigned long long)s_cell[r] << 44) | ((unsigned long long)(uint32_t)s_pid[r] << 13) |
                                    (unsigned long long)r
                              : ~0ull;
            }
            __syncthreads();
            const int P = pow2_ceil(n);
#pragma unroll
            for (int k = 0; k < CAP / WG; ++k) {
                const int r = threadIdx.x + k * WG;
                if (r < P) s_key[r] = kv[k];
            }
            __syncthreads();
            bitonic_lds(s_key, P);
            // final position of each arrival: keys to registers, then positions
            // over the keys, then each arrival placed from registers
#pragma unroll
            for (int k = 0; k < CAP / WG; ++k) {
                const int q = threadIdx.x + k * WG;
                kv[k] = q < n ? s_key[q] : 0ull;
            }
            __syncthreads();
            unsigned short* s_pos = reinterpret_cast<unsigned short*>(s_mem);
#pragma unroll
            for (int k = 0; k < CAP / WG; ++k) {
                const int q = threadIdx.x + k * WG;
                if (q < n) s_pos[(int)(kv[k] & 0x1fffull)] = (unsigned short)q;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < CAP / WG; ++k) {
                const int r = threadIdx.x + k * WG;
                if (r < n) idx[b + s_pos[r]] = svr[k];
            }
        } else {
            for (int r0 = threadIdx.x; r0 < n; r0 += WG * ES_GROUP) {
                int sv[ES_GROUP], kc[ES_GROUP];
#pragma unroll
                for (int j = 0; j < ES_GROUP; ++j) sv[j] = tmp[b + min(r0 + j * WG, n - 1)];
#pragma unroll
                for (int j = 0; j < ES_GROUP; ++j) kc[j] = kcell[sv[j]];
#pragma unroll
                for (int j = 0; j < ES_GROUP; ++j)
                    if (r0 + j * WG < n) idx[b + atomicAdd(&s_hist[kc[j]], 1)] = sv[j];
            }
            __syncthreads();
            for (int c0 = 0; c0 < TILE_CELLS;) {
                if (threadIdx.x == 0) { // extend the round by whole cells up to ROUND particles
                    const int base = s_off[c0];
                    int c1 = c0 + 1;
                    while (c1 < TILE_CELLS && ((c1 + 1 < TILE_CELLS) ? s_off[c1 + 1] : n) - base <= ROUND) ++c1;
                    s_c1 = c1;
                }
                __syncthreads();
                const int c1 = s_c1;
                const int lo = s_off[c0], len = ((c1 < TILE_CELLS) ? s_off[c1] : n) - lo;
                if (len <= ROUND) {
                    const int P = pow2_ceil(len > 0 ? len : 1);
                    for (int q0 = threadIdx.x; q0 < P; q0 += WG * ES_GROUP) {
                        int sv[ES_GROUP], kc[ES_GROUP], pd[ES_GROUP];
                        if (len > 0) { // reads grouped as in the staging above
#pragma unroll
                            for (int j = 0; j < ES_GROUP; ++j) sv[j] = idx[b + lo + min(q0 + j * WG, len - 1)];
#pragma unroll
                            for (int j = 0; j < ES_GROUP; ++j) {
                                kc[j] = kcell[sv[j]];
                                pd[j] = pid[sv[j]];
                            }
                        }
#pragma unroll
                        for (int j = 0; j < ES_GROUP; ++j) {
                            const int q = q0 + j * WG;
                            if (q >= P) continue;
                            unsigned long long kq = ~0ull;
                            if (q < len) {
                                s_sv[q] = sv[j];
                                kq = ((unsigned long long)kc[j] << 44) |
                                     ((unsigned long long)(uint32_t)pd[j] << 13) | (unsigned long long)q;
                            }
                            s_key[q] = kq;
                        }
                    }
                    __syncthreads();
                    bitonic_lds(s_key, P);
                    for (int q = threadIdx.x; q < len; q += WG) idx[b + lo + q] = s_sv[(int)(s_key[q] & 0x1fffull)];
                } else if ((int)threadIdx.x == c0) { // one cell beyond ROUND: in global memory
                    const int glo = b + off, ghi = b + off + cnt;
                    for (int q = glo + 1; q < ghi; ++q) {
                        const int sv = idx[q];
                        const int pv = pid[sv];
                        int k = q - 1;
                        while (k >= glo && pid[idx[k]] > pv) {
                            idx[k + 1] = idx[k];
                            --k;
                        }
                        idx[k + 1] = sv;
                    }
                }
                __syncthreads();
                c0 = c1;
            }
        }
        __syncthreads();
    }
}

// --- P2G -------------------------------------------------------------------

// P2G, src/simulation.cpp:726-798, as a per-tile segmented reduction without
// atomics: one thread per stencil-base cell sums its particles (canonical
// order) into 9 stencil slots x 7 fields in registers; 9 phases then add slot
// (d,e) of every cell into node cell+(d,e) of an 18x18-node LDS region (one-to
// -one within a phase: no conflicts, fixed order); the region is this tile's
// partial, merged per node by k_egrid.
constexpr int P2G_TCELLS = TILE + 1;
constexpr int P2G_LDS_FLOATS = P2G_FIELDS * REGION_NODES + 3 * P2G_TCELLS * P2G_TCELLS;

// The cell (tile row << 4 | column) a P2G thread sums: the two 16-lane rows
// of each 32-lane half of a wave are tile rows r and r + 8, whose partial
// region rows (18 nodes apart per row: offsets 18 r and 18 r + 144) fall in
// disjoint halves of 32 LDS banks, so the 9-phase reduction's adds meet no
// bank conflicts (rows r and r + 1 overlap in two banks).  Which thread sums
// a cell changes no sum.
__device__ __forceinline__ int p2g_cell(int tid) {
    return (((tid >> 5) + ((tid >> 4) & 1) * 8) << 4) | (tid & 15);
}

// A tile's partial region is summed in LDS field-major (field f of node n at
// f * 324 + n: the 9-phase adds meet no bank conflicts) and stored node-major
// by field group, so the grid kernel reads each source node with two vector
// loads: {m, P_x, P_y, H} as a float4 per node, then {F_x, F_y} as a float2
// per node, then P_z per node (output steps only) -- the same 2,268 floats
// (9,072 bytes, 16-byte aligned in a 256-byte aligned buffer) per tile.
static_assert(P2G_FIELDS == 7 && P2G_PZ == 6, "partial groups: 4 + 2 fields, then P_z");
constexpr int PART_F2 = 4 * REGION_NODES; // float offset of the {F_x, F_y} group
constexpr int PART_PZ = 6 * REGION_NODES; // float offset of P_z
__device__ __forceinline__ void clear_region(float* s, int nreg, int tid) {
    float4* s4 = reinterpret_cast<float4*>(s);
    for (int q = tid; q < (nreg >> 2); q += WG) s4[q] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ void store_region(float* __restrict__ dst, const float* s, bool vz, int tid) {
    constexpr int RN = REGION_NODES;
    for (int n = tid; n < RN; n += WG) {
        reinterpret_cast<float4*>(dst)[n] = make_float4(s[n], s[RN + n], s[2 * RN + n], s[3 * RN + n]);
        reinterpret_cast<float2*>(dst + PART_F2)[n] = make_float2(s[4 * RN + n], s[5 * RN + n]);
        if (vz) dst[PART_PZ + n] = s[PART_PZ + n];
    }
}

// One particle's stencil sums into the 9 slots x 7 fields of its cell's
// thread (cell (lx, ly) of the tile, the floor cells' {Gx, Gy, cos theta} in
// s_ter, 17 x 17 per field).  Both P2G paths call it, so a tile's sums do not
// depend on the path that takes it.  VZ adds the vz momentum (the grid's vz
// output, on output steps only).
template <bool VZ>
__device__ __forceinline__ void ep2g_particle(float (&acc)[9][P2G_FIELDS], const float* s_ter, int lx, int ly,
                                              float g, float inv, float ox, float oy, float mp, float hp, float pvx,
                                              float pvy, float pvz) {
    constexpr int TCELLS = P2G_TCELLS;
    // live particles sit in cells [1, N-3] (outflow parks the others), so the
    // stencil base is never clamped and the floor cell is the base plus
    // (offset < 1/2)
    const int ax = ox < 0.5f ? 1 : 0, ay = oy < 0.5f ? 1 : 0;
    const int fl = (lx + ax) * TCELLS + (ly + ay);
    const float fx = (float)ax + ox, fy = (float)ay + oy;
    const float cth = s_ter[2 * TCELLS * TCELLS + fl];
    const float mgc = mp * g * cth * cth; // m g cos^2 theta
    const float vmx = pvx * mp, vmy = pvy * mp, hm = hp * mp, vzm = pvz * mp;
    // body force m b = -m g cos^2 theta grad z
    const float mbx = -mgc * s_ter[fl], mby = -mgc * s_ter[TCELLS * TCELLS + fl];
    // h*A * sigma_00 with h*A = m / rho and sigma_00 = -rho g cos^2 h / 2
    const float ps = -0.5f * mgc * hp;
    float wx[3], wy[3], dwx[3], dwy[3];
    weights3f(fx, inv, wx, dwx);
    weights3f(fy, inv, wy, dwy);
    // node force m b w - ps grad w with w = w_x[d] w_y[e], per axis one
    // factor times the other axis' weight: x (mbx w_x[d] - ps dw_x[d]) w_y[e],
    // y (mby w_y[e] - ps dw_y[e]) w_x[d]
    float frx[3], fry[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        frx[d] = fmaf(mbx, wx[d], -ps * dwx[d]);
        fry[d] = fmaf(mby, wy[d], -ps * dwy[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const float wgt = wx[d] * wy[e];
            // simulation.cpp:780; the centre weight is at least 1/4 (the
            // middle weight of an axis is in [1/2, 3/4])
            if ((d == 1 && e == 1) || wgt > 3e-3f) {
                float* a = acc[d * 3 + e];
                a[0] = fmaf(mp, wgt, a[0]);
                a[1] = fmaf(vmx, wgt, a[1]);
                a[2] = fmaf(vmy, wgt, a[2]);
                a[3] = fmaf(hm, wgt, a[3]);
                a[4] = fmaf(frx[d], wy[e], a[4]);
                a[5] = fmaf(fry[e], wx[d], a[5]);
                if (VZ) a[P2G_PZ] = fmaf(vzm, wgt, a[P2G_PZ]);
            }
        }
    }
}
// s_acc: P2G_FIELDS x REGION_NODES floats, s_ter: 3 x 17 x 17 floats
__device__ __forceinline__ void ep2g_tiles(ECtl* c, const Slot* __restrict__ slots, PB32 A, PB32 B,
                                           const int* __restrict__ idx, const int* __restrict__ ptiles,
                                           const int* __restrict__ cell_begin, const float* __restrict__ tgx,
                                           const float* __restrict__ tgy, const float* __restrict__ tcth,
                                           float* __restrict__ tpart, const uint8_t* __restrict__ only,
                                           float* __restrict__ s_acc, float* __restrict__ s_ter) {
    constexpr int TCELLS = P2G_TCELLS;
    const PB32 P = ecur(c, A, B);
    const int nt = c->n_ptiles, nx = c->nx, ny = c->ny, nty = c->nty, ntd = c->ntd, live = c->nslots * c->ntd;
    const float inv = (float)c->inv_dx;
    const int cid = p2g_cell(threadIdx.x);
    const int lx = cid >> 4, ly = cid & 15;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        if (!only[w]) continue; // the tiles the small-tile pass left
        // this thread's particle range in the sorted index
        const int* cb = cell_begin + (size_t)w * CELL_STRIDE;
        const int r0 = cb[cid], r1 = cb[cid + 1];
        // this thread's first particle's loads depend on its range only: they
        // are issued before the tile and slot checks and the LDS staging below
        // (every listed range lies inside the sorted index), and overlap them
        float nox = 0.f, noy = 0.f, nmp = 0.f, nhp = 0.f, nvx = 0.f, nvy = 0.f, nvz = 0.f;
        int sn0 = 0;
        if (r0 < r1) {
            sn0 = idx[r0];
            const float2 po = P.o[sn0], pv = P.v[sn0];
            nox = po.x; noy = po.y; nmp = P.m[sn0]; nhp = P.h[sn0];
            nvx = pv.x; nvy = pv.y;
        }
        const int t = ptiles[w];
        if (t >= live) continue;
        const int slot = t / ntd, local = t - slot * ntd;
        if (!slot_active(slots[slot].status)) continue;
        const float g = (float)slots[slot].gravity;
        // P_z only feeds the grid's vz output, written on output steps
        const bool need_vz = slots[slot].out_step != 0;
        if (need_vz && r0 < r1) nvz = P.vz[sn0];
        const int tx = local / nty, ty = local - tx * nty;
        // the region's fields; P_z (the last) on output steps only
        const int nreg = (need_vz ? P2G_FIELDS : P2G_PZ) * REGION_NODES;
        clear_region(s_acc, nreg, threadIdx.x);
        // the tile's floor cells (base + {0,1}) hold {Gx, Gy, cos theta}: staged
        // once, every load of a thread issued before its first LDS write
        constexpr int TPT = (TCELLS * TCELLS + WG - 1) / WG;
        float tv[TPT][3];
#pragma unroll
        for (int i = 0; i < TPT; ++i) {
            const int q = threadIdx.x + i * WG;
            const int a = q / TCELLS, b = q - a * TCELLS;
            const int gi = tx * TILE + a, gj = ty * TILE + b;
            tv[i][0] = tv[i][1] = tv[i][2] = 0.f;
            if (q < TCELLS * TCELLS && gi < nx && gj < ny) {
                const size_t k = (size_t)gi * ny + gj;
                tv[i][0] = tgx[k];
                tv[i][1] = tgy[k];
                tv[i][2] = tcth[k];
            }
        }
#pragma unroll
        for (int i = 0; i < TPT; ++i) {
            const int q = threadIdx.x + i * WG;
            if (q < TCELLS * TCELLS) {
                s_ter[q] = tv[i][0];
                s_ter[TCELLS * TCELLS + q] = tv[i][1];
                s_ter[2 * TCELLS * TCELLS + q] = tv[i][2];
            }
        }
        __syncthreads();
        float acc[9][P2G_FIELDS];
#pragma unroll
        for (int k = 0; k < 9; ++k)
#pragma unroll
            for (int f = 0; f < P2G_FIELDS; ++f) acc[k][f] = 0.f;
        // the next particle's loads are issued before the current one is
        // summed, and the index of the one after it a particle further ahead,
        // so no iteration waits for an index load; the vz momentum (grid vz
        // output only) is summed on output steps only
        auto sum_particles = [&](auto vz_tag) {
        constexpr bool VZ = decltype(vz_tag)::value;
        int sn_next = r0 + 1 < r1 ? idx[r0 + 1] : 0;
        for (int r = r0; r < r1; ++r) {
            const float ox = nox, oy = noy, mp = nmp, hp = nhp, pvx = nvx, pvy = nvy, pvz = nvz;
            if (r + 1 < r1) {
                const int sn = sn_next;
                const float2 po = P.o[sn], pv = P.v[sn];
                nox = po.x; noy = po.y; nmp = P.m[sn]; nhp = P.h[sn];
                nvx = pv.x; nvy = pv.y;
                if (VZ) nvz = P.vz[sn];
                if (r + 2 < r1) sn_next = idx[r + 2];
            }
            ep2g_particle<VZ>(acc, s_ter, lx, ly, g, inv, ox, oy, mp, hp, pvx, pvy, pvz);
        }
        };
        if (need_vz)
            sum_particles(std::true_type{});
        else
            sum_particles(std::false_type{});
        __syncthreads();
#pragma unroll
        for (int ph = 0; ph < 9; ++ph) {
            const int node = (lx + ph / 3) * REGION + (ly + ph % 3);
#pragma unroll
            for (int f = 0; f < P2G_FIELDS; ++f)
                if (f < P2G_PZ || need_vz) s_acc[f * REGION_NODES + node] += acc[ph][f];
            __syncthreads();
        }
        float* dst = tpart + (size_t)w * (P2G_FIELDS * REGION_NODES);
        store_region(dst, s_acc, need_vz, threadIdx.x);
        __syncthreads();
    }
}

// The sort + P2G kernel's LDS region (beside the histogram arrays) for an
// in-LDS sort capacity of cap particles: the sort's three arrays (9 B per
// particle), or the P2G's sorted storage indices, partial region and terrain.
constexpr int esp_lds_bytes(int cap) {
    return cap * 9 > cap * 4 + P2G_LDS_FLOATS * 4 ? cap * 9 : cap * 4 + P2G_LDS_FLOATS * 4;
}
constexpr int P2G_TER_BYTES = 3 * P2G_TCELLS * P2G_TCELLS * 4;
// particle fields a small tile's P2G stages in LDS: offset, velocity, mass,
// height (24 B), and vz on output steps (28 B)
constexpr int ESP_STAGE_B = 24, ESP_STAGE_B_VZ = 28;

// Small tiles (at most CAP particles, no cell above ES_RANK): sorted and
// summed in one pass per tile -- placed by rank into the tile's sorted order
// in LDS, then P2G over that order (the sorted index is written only on
// re-sort steps, whose G2P walks it).  Other tiles are marked in defer[w].
template <int CAP>
__device__ __forceinline__ void esp_small_tiles(ECtl* c, const Slot* __restrict__ slots, PB32 A, PB32 B,
                                                const uint8_t* __restrict__ kcell, const int* __restrict__ ptiles,
                                                const int2* __restrict__ pinfo, int* __restrict__ tile_count,
                                                int* __restrict__ cursor, const int* __restrict__ tmp,
                                                int* __restrict__ idx, uint8_t* __restrict__ defer,
                                                const float* __restrict__ tgx, const float* __restrict__ tgy,
                                                const float* __restrict__ tcth, float* __restrict__ tpart,
                                                const ESortLds& L) {
    static_assert(CAP % (WG * ES_GROUP) == 0 && CAP <= 8192, "sort capacity");
    constexpr int TCELLS = P2G_TCELLS;
    int* s_hist = L.hist;
    int* s_off = L.off;
    int* s_w = L.w4;
    int& s_cmax = *L.cmax;
    unsigned char* s_mem = L.mem;
    // the LDS region: (a) creation index by arrival, creation index in cell
    // order and cell of the tile's particles; (d) P2G: the tile's sorted
    // storage indices (over (a)'s first array), the 18x18-node partial region
    // and, at the region's end, the floor cells' terrain.  A staged tile's
    // P2G keeps its particle fields (SoA, sorted order) below the terrain,
    // over the sorted indices, and its partial region over the fields once
    // they are summed.
    constexpr int LDS_BYTES = esp_lds_bytes(CAP);
    constexpr int STAGE_BYTES = LDS_BYTES - P2G_TER_BYTES;
    constexpr int ESK = (STAGE_BYTES / ESP_STAGE_B + WG - 1) / WG; // staged particles per thread
    static_assert(CAP * 4 + P2G_FIELDS * REGION_NODES * 4 <= STAGE_BYTES, "P2G LDS layout");
    int* s_pid = reinterpret_cast<int*>(s_mem);                                    // (a) CAP x i32
    int* s_pidc = reinterpret_cast<int*>(s_mem + CAP * 4);                         // (a) CAP x i32
    unsigned char* s_cell = s_mem + CAP * 8;                                       // (a) CAP x u8
    int* s_sorted = reinterpret_cast<int*>(s_mem);                                 // (d) CAP x i32
    float* s_acc = reinterpret_cast<float*>(s_mem + CAP * 4);                      // (d) 7 x 324
    float* s_ter = reinterpret_cast<float*>(s_mem + STAGE_BYTES);                  // (d) 3 x 17 x 17
    const PB32 P = ecur(c, A, B);
    const int32_t* __restrict__ pid = P.pid;
    const int nt = c->n_ptiles, nx = c->nx, ny = c->ny, nty = c->nty, ntd = c->ntd, live = c->nslots * c->ntd;
    const bool reo = c->reorder != 0, compact = c->compact != 0;
    const float inv = (float)c->inv_dx;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        // the thread index through an empty asm, once per tile: what derives
        // from it is recomputed per tile instead of held in registers across
        // the sort and P2G phases
        const int tid = tid_opaque();
        // the tile, its begin and its count, by list position: read-only here,
        // so read together through the scalar cache
        const int t = ptiles[w];
        const int2 nb = pinfo[w];
        const int n = nb.y, b = nb.x; // n >= 1: listed tiles hold particles
        if (n > CAP && t < live) { // left to the sort in global memory
            if (tid == 0) defer[w] = 1;
            continue;
        }
        // the first group of bucket reads and their key and creation-index
        // reads touch no LDS: issued before the barrier, they overlap it
        int sv0[ES_GROUP], kc0[ES_GROUP], pd0[ES_GROUP];
#pragma unroll
        for (int j = 0; j < ES_GROUP; ++j) sv0[j] = tmp[b + min((int)tid + j * WG, n - 1)];
#pragma unroll
        for (int j = 0; j < ES_GROUP; ++j) {
            kc0[j] = kcell[sv0[j]];
            pd0[j] = pid[sv0[j]];
        }
        if (tid == 0) s_cmax = 0;
        s_hist[tid] = 0;
        __syncthreads();
        if (t >= live) { // parking tile
            for (int r = tid; r < n; r += WG) idx[b + r] = tmp[b + r];
            if (tid == 0) {
                defer[w] = 0;
                tile_count[t] = 0;
                cursor[t] = 0;
            }
            __syncthreads();
            continue;
        }
        // storage index of each arrival this thread stages: the final write
        // places it from registers instead of re-reading tmp
        int svr[CAP / WG];
        {
        // groups of ES_GROUP arrivals per thread: the bucket reads of a group
        // are issued together, then their key and creation-index reads
        // (clamped indices keep every read in the tile); the next group's
        // bucket reads are issued before this group's keys are waited for
        int svn[ES_GROUP];
#pragma unroll
        for (int j = 0; j < ES_GROUP; ++j)
            svn[j] = ES_GROUP * WG < n ? tmp[b + min((int)tid + (ES_GROUP + j) * WG, n - 1)] : 0;
#pragma unroll
        for (int k0 = 0; k0 < CAP / WG; k0 += ES_GROUP) {
            if (k0 * WG < n) {
                int sv[ES_GROUP], kc[ES_GROUP], pd[ES_GROUP];
                if (k0 == 0) {
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) {
                        sv[j] = sv0[j];
                        kc[j] = kc0[j];
                        pd[j] = pd0[j];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) sv[j] = svn[j];
                    if (k0 + ES_GROUP < CAP / WG && (k0 + ES_GROUP) * WG < n) {
#pragma unroll
                        for (int j = 0; j < ES_GROUP; ++j)
                            svn[j] = tmp[b + min((int)tid + (k0 + ES_GROUP + j) * WG, n - 1)];
                    }
#pragma unroll
                    for (int j = 0; j < ES_GROUP; ++j) {
                        kc[j] = kcell[sv[j]];
                        pd[j] = pid[sv[j]];
                    }
                }
#pragma unroll
                for (int j = 0; j < ES_GROUP; ++j) {
                    const int r = tid + (k0 + j) * WG;
                    svr[k0 + j] = sv[j];
                    if (r < n) {
                        const int cc = kc[j];
                        s_cell[r] = (unsigned char)cc;
                        s_pid[r] = pd[j];
                        atomicAdd(&s_hist[cc], 1);
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < ES_GROUP; ++j) svr[k0 + j] = 0;
            }
        }
        }
        __syncthreads();
        const int cnt = s_hist[tid];
        if (cnt > ES_RANK) atomicMax(&s_cmax, cnt);
        {
            const int lane = tid & 63, wv = tid >> 6;
            int inc = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                int u = __shfl_up(inc, o, 64);
                if (lane >= o) inc += u;
            }
            if (lane == 63) s_w[wv] = inc;
            __syncthreads();
            int pre = 0;
            for (int q = 0; q < wv; ++q) pre += s_w[q];
            s_off[tid] = pre + inc - cnt;
            if (tid == 0) {
                tile_count[t] = 0;
                cursor[t] = 0;
            }
        }
        __syncthreads();
        const int off = s_off[tid];
        s_hist[tid] = off; // cursor
        const bool dense = s_cmax > ES_RANK;
        __syncthreads();
        if (dense) { // a cell above ES_RANK particles: left to the bitonic sort
            if (tid == 0) defer[w] = 1;
            continue;
        }
        if (tid == 0) defer[w] = 0;
        // each arrival's creation index, also kept in registers: the sorted
        // order is written over s_pid below
        int pdr[CAP / WG];
#pragma unroll
        for (int k = 0; k < CAP / WG; ++k) {
            const int r = tid + k * WG;
            pdr[k] = 0;
            if (r < n) {
                pdr[k] = s_pid[r];
                s_pidc[atomicAdd(&s_hist[s_cell[r]], 1)] = pdr[k];
            }
        }
        __syncthreads();
        // s_pidc[s_off[c] .. s_hist[c]) now holds cell c's creation indices;
        // each arrival is placed from its registers at the cell start plus
        // its rank (creation indices are unique within the tile's slot): into
        // the tile's sorted order in LDS for P2G, and into the sorted index on
        // a re-sort step (whose G2P walks it)
#pragma unroll
        for (int k = 0; k < CAP / WG; ++k) {
            const int r = tid + k * WG;
            if (r < n) {
                const int cr = s_cell[r], pr = pdr[k];
                const int lo = s_off[cr], hi = s_hist[cr];
                int rank = 0;
                for (int q = lo; q < hi; ++q) rank += s_pidc[q] < pr ? 1 : 0;
                s_sorted[lo + rank] = svr[k];
                if (reo) idx[b + lo + rank] = svr[k];
            }
        }
        __syncthreads();
        // ---- P2G of the tile, src/simulation.cpp:726-798: one thread per
        // stencil-base cell sums its particles (canonical order) into 9 stencil
        // slots x 7 fields in registers; 9 phases then add slot (d,e) of every
        // cell into node cell+(d,e) of an 18x18-node LDS region (one-to-one
        // within a phase: no conflicts, fixed order); the region is this
        // tile's partial, merged per node by k_egrid.  A re-sort pass has no
        // P2G; neither has a slot that is not running.
        if (compact) continue;
        const int slot = t / ntd, local = t - slot * ntd;
        if (!slot_active(slots[slot].status)) continue;
        const float g = (float)slots[slot].gravity;
        // P_z only feeds the grid's vz output, written on output steps
        const bool need_vz = slots[slot].out_step != 0;
        // this thread's cell's particles: positions [r0, r1) of the tile's
        // sorted order
        const int pc = p2g_cell(tid);
        const int lx = pc >> 4, ly = pc & 15;
        const int r0 = s_off[pc], r1 = s_hist[pc]; // s_hist: the cell's end after the placement
        const int tx = local / nty, ty = local - tx * nty;
        // a tile whose particle fields fit the LDS below the terrain has them
        // staged there in sorted order: consecutive positions on consecutive
        // lanes, all of a thread's reads issued together, so the summing loop
        // waits on no global read.  A larger tile's threads read their cell's
        // particles from global memory as they go, one particle ahead.  The
        // two forms are separate code paths, so neither holds the other's
        // registers.
        auto p2g_tile = [&](auto staged_tag) {
            constexpr bool STAGED = decltype(staged_tag)::value;
            float2 so[STAGED ? ESK : 1], sw[STAGED ? ESK : 1];
            float sm[STAGED ? ESK : 1], sh[STAGED ? ESK : 1], sz[STAGED ? ESK : 1];
            float nox = 0.f, noy = 0.f, nmp = 0.f, nhp = 0.f, nvx = 0.f, nvy = 0.f, nvz = 0.f;
            if constexpr (STAGED) {
#pragma unroll
                for (int k = 0; k < ESK; ++k) {
                    const int q = tid + k * WG;
                    so[k] = sw[k] = make_float2(0.f, 0.f);
                    sm[k] = sh[k] = sz[k] = 0.f;
                    if (q < n) {
                        const int sn = s_sorted[q];
                        so[k] = P.o[sn];
                        sw[k] = P.v[sn];
                        sm[k] = P.m[sn];
                        sh[k] = P.h[sn];
                        if (need_vz) sz[k] = P.vz[sn];
                    }
                }
            } else if (r0 < r1) { // the first particle's loads
                const int sn0 = s_sorted[r0];
                const float2 po = P.o[sn0], pv = P.v[sn0];
                nox = po.x; noy = po.y; nmp = P.m[sn0]; nhp = P.h[sn0];
                nvx = pv.x; nvy = pv.y;
                if (need_vz) nvz = P.vz[sn0];
            }
            // the tile's floor cells (base + {0,1}) hold {Gx, Gy, cos theta}:
            // staged once, every load of a thread issued before its first LDS
            // write
            constexpr int TPT = (TCELLS * TCELLS + WG - 1) / WG;
            float tv[TPT][3];
#pragma unroll
            for (int i = 0; i < TPT; ++i) {
                const int q = tid + i * WG;
                const int qa = q / TCELLS, qb = q - qa * TCELLS;
                const int gi = tx * TILE + qa, gj = ty * TILE + qb;
                tv[i][0] = tv[i][1] = tv[i][2] = 0.f;
                if (q < TCELLS * TCELLS && gi < nx && gj < ny) {
                    const size_t k = (size_t)gi * ny + gj;
                    tv[i][0] = tgx[k];
                    tv[i][1] = tgy[k];
                    tv[i][2] = tcth[k];
                }
            }
            // the region's fields; P_z (the last) on output steps only.  A
            // staged tile's region lies over its fields and is cleared after
            // the sums.
            const int nreg = (need_vz ? P2G_FIELDS : P2G_PZ) * REGION_NODES;
            float* const s_reg = STAGED ? reinterpret_cast<float*>(s_mem) : s_acc;
            if (!STAGED)
                clear_region(s_acc, nreg, tid);
#pragma unroll
            for (int i = 0; i < TPT; ++i) {
                const int q = tid + i * WG;
                if (q < TCELLS * TCELLS) {
                    s_ter[q] = tv[i][0];
                    s_ter[TCELLS * TCELLS + q] = tv[i][1];
                    s_ter[2 * TCELLS * TCELLS + q] = tv[i][2];
                }
            }
            float2* const s_o2 = reinterpret_cast<float2*>(s_mem);
            float2* const s_v2 = s_o2 + n;
            float* const s_m = reinterpret_cast<float*>(s_v2 + n);
            float* const s_h = s_m + n;
            float* const s_z = s_h + n;
            if constexpr (STAGED) {
                __syncthreads(); // every thread has read the sorted indices the fields go over
#pragma unroll
                for (int k = 0; k < ESK; ++k) {
                    const int q = tid + k * WG;
                    if (q < n) {
                        s_o2[q] = so[k];
                        s_v2[q] = sw[k];
                        s_m[q] = sm[k];
                        s_h[q] = sh[k];
                        if (need_vz) s_z[q] = sz[k];
                    }
                }
            }
            __syncthreads();
            float acc[9][P2G_FIELDS];
#pragma unroll
            for (int k = 0; k < 9; ++k)
#pragma unroll
                for (int f = 0; f < P2G_FIELDS; ++f) acc[k][f] = 0.f;
            // the vz momentum (grid vz output only) is summed on output steps
            // only.  Unstaged: the next particle's loads are issued before the
            // current one is summed, and the index of the one after it a
            // particle further ahead, so no iteration waits for an index load.
            auto sum_particles = [&](auto vz_tag) {
                constexpr bool VZ = decltype(vz_tag)::value;
                int sn_next = !STAGED && r0 + 1 < r1 ? s_sorted[r0 + 1] : 0;
                for (int r = r0; r < r1; ++r) {
                    float ox, oy, mp, hp, pvx, pvy, pvz;
                    if constexpr (STAGED) {
                        const float2 po = s_o2[r], pv = s_v2[r];
                        ox = po.x; oy = po.y; mp = s_m[r]; hp = s_h[r];
                        pvx = pv.x; pvy = pv.y;
                        pvz = VZ ? s_z[r] : 0.f;
                    } else {
                        ox = nox; oy = noy; mp = nmp; hp = nhp; pvx = nvx; pvy = nvy; pvz = nvz;
                        if (r + 1 < r1) {
                            const int sn = sn_next;
                            const float2 po = P.o[sn], pv = P.v[sn];
                            nox = po.x; noy = po.y; nmp = P.m[sn]; nhp = P.h[sn];
                            nvx = pv.x; nvy = pv.y;
                            if (VZ) nvz = P.vz[sn];
                            if (r + 2 < r1) sn_next = s_sorted[r + 2];
                        }
                    }
                    ep2g_particle<VZ>(acc, s_ter, lx, ly, g, inv, ox, oy, mp, hp, pvx, pvy, pvz);
                }
            };
            if (need_vz)
                sum_particles(std::true_type{});
            else
                sum_particles(std::false_type{});
            if constexpr (STAGED) {
                __syncthreads(); // the fields are summed: the region goes over them
                clear_region(s_reg, nreg, tid);
            }
            __syncthreads();
#pragma unroll
            for (int ph = 0; ph < 9; ++ph) {
                const int node = (lx + ph / 3) * REGION + (ly + ph % 3);
#pragma unroll
                for (int f = 0; f < P2G_FIELDS; ++f)
                    if (f < P2G_PZ || need_vz) s_reg[f * REGION_NODES + node] += acc[ph][f];
                __syncthreads();
            }
            float* dst = tpart + (size_t)w * (P2G_FIELDS * REGION_NODES);
            store_region(dst, s_reg, need_vz, tid);
            __syncthreads();
        };
        if (n * (need_vz ? ESP_STAGE_B_VZ : ESP_STAGE_B) <= STAGE_BYTES)
            p2g_tile(std::true_type{});
        else
            p2g_tile(std::false_type{});
    }
}

// Sort and P2G in one launch, in three passes over the workgroup's tiles (list
// positions blockIdx.x + k * gridDim.x): the small tiles, sorted and summed
// from LDS; then the tiles they left, sorted in LDS rounds / global memory
// into the sorted index; then P2G of those from the index.  No kernel
// boundary between the sort and P2G, one LDS region for all passes, and each
// pass a loop of its own, so none holds another's registers.
template <int CAP>
__global__ void __launch_bounds__(WG) k_esp(ECtl* c, const Slot* __restrict__ slots, PB32 A, PB32 B,
                                            const uint8_t* __restrict__ kcell, const int* __restrict__ ptiles,
                                            const int2* __restrict__ pinfo, int* __restrict__ tile_count,
                                            int* __restrict__ cursor, const int* __restrict__ tmp,
                                            int* __restrict__ idx, int* __restrict__ cell_begin,
                                            uint8_t* __restrict__ defer, const float* __restrict__ tgx,
                                            const float* __restrict__ tgy, const float* __restrict__ tcth,
                                            float* __restrict__ tpart) {
    if (c->stop) return;
    __shared__ int s_hist[TILE_CELLS];
    __shared__ int s_off[TILE_CELLS];
    __shared__ int s_w[4];
    __shared__ int s_c1, s_cmax;
    __shared__ __align__(16) unsigned char s_mem[esp_lds_bytes(CAP)];
    const ESortLds L{s_hist, s_off, s_w, &s_c1, &s_cmax, s_mem};
    esp_small_tiles<CAP>(c, slots, A, B, kcell, ptiles, pinfo, tile_count, cursor, tmp, idx, defer, tgx, tgy, tcth,
                         tpart, L);
    // the pass's marks and sorted tiles are visible to the whole workgroup
    __syncthreads();
    esort_tiles<CAP>(c, A, B, kcell, ptiles, pinfo, tile_count, cursor, tmp, idx, cell_begin, defer, L);
    if (c->compact) return;
    __syncthreads();
    float* s_p2g = reinterpret_cast<float*>(s_mem);
    ep2g_tiles(c, slots, A, B, idx, ptiles, cell_begin, tgx, tgy, tcth, tpart, defer, s_p2g,
               s_p2g + P2G_FIELDS * REGION_NODES);
}

// --- grid ------------------------------------------------------------------

// computeGridVelocity (src/simulation.cpp:800-875) for one grid tile per
// workgroup, one thread per node, after merging the node's partial sums from
// the particle tiles of the same slot that reach it (own, left, below,
// diagonal; fixed order).
// The grid velocity / FLIP field {vx, vy, flipx, flipy} it writes lives for
// one step.  A single run keeps it dense over the terrain (slot * nn + i*ny +
// j).  A batch (POOL) keeps it in a pool of the step's grid tiles only: tile
// w of the grid-tile list holds its 256 nodes at w*256 + (i&15)*16 + (j&15),
// and gtile_pos maps a tile id to its list position, so the batch's memory
// grows with its active tiles instead of slots x terrain.  A batch's max maps
// h_max, v_max2 live in a second pool, of the tiles its slots have touched:
// hm_map maps a tile id to its pool tile (256 nodes, same local order).
template <bool POOL>
__global__ void __launch_bounds__(WG) k_egrid(ECtl* c, const Slot* __restrict__ slots,
                                              const int* __restrict__ gtiles, const int* __restrict__ pstamp,
                                              const int* __restrict__ tile_pos, const float* __restrict__ tpart,
                                              const float* __restrict__ tgx, const float* __restrict__ tgy,
                                              const float* __restrict__ tcth, float* __restrict__ h_max,
                                              float* __restrict__ v_max2, float4* __restrict__ vf,
                                              float* __restrict__ out_m, float* __restrict__ out_h,
                                              float* __restrict__ out_vz, const int* __restrict__ hm_map) {
    // multiply-adds fused per source expression, so the two layouts (which
    // differ only in addressing) round the same way
#pragma clang fp contract(on)
    if (c->stop || c->compact) return;
    const int nt = c->n_gtiles, ep = c->epoch, nx = c->nx, ny = c->ny, nty = c->nty, ntd = c->ntd;
    const size_t nn = (size_t)nx * ny;
    const int a = threadIdx.x >> 4, b = threadIdx.x & 15;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = gtiles[w];
        const int slot = t / ntd, local = t - slot * ntd;
        const Slot& S = slots[slot];
        if (!slot_active(S.status)) continue;
        const int tx = local / nty, ty = local - tx * nty;
        const int i = tx * TILE + a, j = ty * TILE + b;
        if (i >= nx || j >= ny) continue;
        float acc[P2G_FIELDS];
#pragma unroll
        for (int f = 0; f < P2G_FIELDS; ++f) acc[f] = 0.f;
        // sources in fixed order: own, left, below, diagonal tile.  The tile
        // ids depend only on the tile (a missing neighbour reads the own
        // tile's entries, unused), so their stamps and list positions are
        // read once per workgroup, together (a stale position is never used);
        // then the partial values of the stamped sources that reach this node
        // are read before any is summed (sums in source order).  The node's
        // terrain and max maps do not depend on them and are read alongside.
        const int src[4] = {t, tx > 0 ? t - nty : t, ty > 0 ? t - 1 : t, tx > 0 && ty > 0 ? t - nty - 1 : t};
        const bool hx = tx > 0 && a < 2, hy = ty > 0 && b < 2;
        const bool has[4] = {true, hx, hy, hx && hy};
        const int rn[4] = {a * REGION + b, (a + TILE) * REGION + b, a * REGION + (b + TILE),
                           (a + TILE) * REGION + (b + TILE)};
        const size_t kt = (size_t)i * ny + j; // terrain node
        const size_t k = (size_t)slot * nn + kt;
        const size_t km = POOL ? (size_t)hm_map[t] * TILE_CELLS + threadIdx.x : k;
        const float Gx = tgx[kt], Gy = tgy[kt], cth = tcth[kt];
        const float hm_old = h_max[km], vm_old = v_max2[km];
        bool ok[4];
        int pos[4];
        float val[4][P2G_FIELDS];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int st = pstamp[src[q]], ps = tile_pos[src[q]];
            ok[q] = has[q] && st == ep;
            pos[q] = ps;
        }
        // P_z (the last field) is stored on output steps only
        const bool pz = S.out_step != 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!ok[q]) continue;
            const float* pp = tpart + (size_t)pos[q] * (P2G_FIELDS * REGION_NODES);
            const float4 m4 = reinterpret_cast<const float4*>(pp)[rn[q]];
            const float2 f2 = reinterpret_cast<const float2*>(pp + PART_F2)[rn[q]];
            val[q][0] = m4.x;
            val[q][1] = m4.y;
            val[q][2] = m4.z;
            val[q][3] = m4.w;
            val[q][4] = f2.x;
            val[q][5] = f2.y;
            val[q][P2G_PZ] = pz ? pp[PART_PZ + rn[q]] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!ok[q]) continue;
#pragma unroll
            for (int f = 0; f < P2G_FIELDS; ++f) acc[f] += val[q][f];
        }
        const float g = (float)S.gravity, dt = (float)S.dt;
        const float am = acc[0];
        float nvx = 0.f, nvy = 0.f, flx = 0.f, fly = 0.f, hh = 0.f, vz = 0.f;
        if (am > 1e-6f) {
            const float rm = 1.0f / am;
            const float vx = acc[1] * rm, vy = acc[2] * rm;
            vz = acc[P2G_PZ] * rm;
            hh = acc[3] * rm;
            const float vtx = vx + dt * (acc[4] * rm), vty = vy + dt * (acc[5] * rm);
            const float vsq = vtx * vtx + vty * vty;
            if (vsq > 1e-16f) {
                const float v_z = vtx * Gx + vty * Gy;
                const float v3sq = vsq + v_z * v_z;
                const float vn = sqrtf(v3sq);
                float af;
                if (S.fric == 0) {
                    af = (float)S.mu * g * cth;
                } else if (S.fric == 1) {
                    af = (float)S.mu * g * cth + g * v3sq / ((float)S.xi * hh * cth);
                } else {
                    const float I = 5.f * (float)S.d * vn / (2.f * hh * cth * cth * sqrtf(g * hh * (float)S.phi));
                    const float muI = (float)S.mu1 + ((float)S.mu2 - (float)S.mu1) / (1.f + (float)S.I0 / I);
                    af = muI * g * cth;
                }
                const float dv = dt * af;
                if (dv < vn) {
                    const float f = 1.f - dv / vn;
                    nvx = vtx * f;
                    nvy = vty * f;
                }
            }
            flx = nvx - vx;
            fly = nvy - vy;
            if (hm_old < hh) h_max[km] = hh;
            const float v2 = nvx * nvx + nvy * nvy;
            if (vm_old < v2) v_max2[km] = v2;
        }
        vf[POOL ? (size_t)w * TILE_CELLS + threadIdx.x : k] = make_float4(nvx, nvy, flx, fly);
        if (out_m && S.out_step) {
            out_m[k] = am;
            out_h[k] = hh;
            out_vz[k] = vz;
        }
    }
}

// --- G2P -------------------------------------------------------------------

// G2P (src/simulation.cpp:877-940) for running slots; priming / finished
// slots are carried unchanged.  On a re-sort step writes the particles to the
// other buffer in canonical order, otherwise updates them in place (the
// constant mass and creation index are not rewritten).  Also writes the next
// keys and tile counts (keys indexed like the storage) and the per-slot max
// |v|^2.  The particle elevation z (simulation.cpp:930) is read by nothing but
// the particle output of a frame (save_function.cpp:58): it is interpolated on
// the slot's output steps only (frames, the prime, every step with keep_grid)
// and carried unchanged otherwise.
// The single-run form with 32-bit offsets fits 7 waves per SIMD (71 VGPRs, no
// spills; 6 waves unhinted): G2P 140 -> 137 us at 8M particles.  The others
// would spill at 7.
template <bool POOL, bool O32>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(POOL || !O32 ? 1 : 7))) k_eg2p(ECtl* c, const Slot* __restrict__ slots, PB32 A, PB32 B,
                                             const int* __restrict__ idx, const float4* __restrict__ vf,
                                             const int* __restrict__ gtile_pos,
                                             const float* __restrict__ tz, const float* __restrict__ tgx,
                                             const float* __restrict__ tgy, int* __restrict__ key,
                                             uint8_t* __restrict__ kcell, int* __restrict__ tile_count,
                                             unsigned long long* __restrict__ vmx) {
    // multiply-adds fused per source expression, so the two layouts (which
    // differ only in addressing) round the same way
#pragma clang fp contract(on)
    if (c->stop) return;
    const PB32 P = ecur(c, A, B);
    const bool reo = c->reorder != 0;
    const PB32 Q = reo ? enxt(c, A, B) : P;
    // A re-sort step walks the particles in canonical order (idx) and writes
    // them to the other buffer in that order.  Other steps update storage in
    // place in storage order: every stored particle is keyed then (release
    // compacts at once), storage was canonical at the last re-sort, and no
    // result depends on the walk order (the sort orders each tile by cell and
    // creation index).
    const int nwork = reo ? c->n_sorted : c->n_store;
    const int r = xcd_block((nwork + WG - 1) / WG) * WG + threadIdx.x;
    const bool in = r < nwork;
    // A single run (dense layout) has one slot: its status, dt and FLIP ratio
    // are read once per workgroup, so no particle waits for a per-particle
    // slot lookup before its stencil reads.
    int st0 = 0, zo0 = 0, vd0 = 0;
    float dt0 = 0.f, fr0 = 0.f;
    if (!POOL) {
        st0 = slots[0].status;
        zo0 = slots[0].out_step;
        vd0 = slots[0].vz_due;
        dt0 = (float)slots[0].dt;
        fr0 = (float)slots[0].flip_ratio;
    }
    const EGeo geo = egeo(c);
    const bool compact = c->compact != 0;
    const float inv = (float)c->inv_dx;
    // node and particle indices: 32-bit when every array's byte offsets fit
    using IX = typename std::conditional<O32, unsigned, size_t>::type;
    int t = 0, kk = -1, sl = 0, o = 0, cb = 0;
    float v2 = 0.f;
    bool pub = false;
    if (in) {
        const int nx = geo.nx, ny = geo.ny;
        const IX nn = (IX)nx * (IX)ny;
        const int s = reo ? el<O32>(idx, r) : r;
        o = r; // storage index written
        // a batch reads the particle's slot; a single run has one slot, and a
        // particle is parked exactly when its cell lies outside [1, N-3]
        // (admission parks those, and G2P parks a particle when it moves
        // there and then never moves it), so its stored flag is not read
        const int raw = POOL ? el<O32>(P.slot, s) : 0;
        sl = POOL ? raw & SLOT_MASK : 0;
        int cell = el<O32>(P.cell, s);
        bool out = POOL ? (raw & SLOT_OUT) != 0 : cell_out(geo, cell >> 16, cell & 0xffff);
        bool parked = false; // parked by this step's move
        const int st = POOL ? slots[sl].status : st0;
        const float2 po = el<O32>(P.o, s), pv = el<O32>(P.v, s);
        float ox = po.x, oy = po.y, vxp = pv.x, vyp = pv.y, hp = el<O32>(P.h, s);
        float vzp = 0.f, zp = 0.f; // recomputed by a running particle, carried otherwise
        bool zset = false;         // zp holds this step's elevation
        bool vzset = false;        // vzp holds this step's vz
        if (st == SLOT_RUNNING && !compact && !out) {
            const float dt = POOL ? (float)slots[sl].dt : dt0, fr = POOL ? (float)slots[sl].flip_ratio : fr0;
            const int ci = cell >> 16, cj = cell & 0xffff;
            // vz = v . grad z at the particle's cell, on the steps that need it
            // (a batch refreshes it every step)
            const bool vzl = POOL || vd0 != 0;
            float Gx = 0.f, Gy = 0.f;
            if (vzl) {
                const IX fc = (IX)ci * (IX)ny + (IX)cj;
                Gx = el<O32>(tgx, fc);
                Gy = el<O32>(tgy, fc);
            }
            const int ib = ebase(ci, ox, nx - 3), jb = ebase(cj, oy, ny - 3);
            const float fx = (float)(ci - ib) + ox, fy = (float)(cj - jb) + oy;
            float wx[3], wy[3], dwx[3], dwy[3];
            weights3f(fx, inv, wx, dwx);
            weights3f(fy, inv, wy, dwy);
            float picx = 0.f, picy = 0.f, flx = 0.f, fly = 0.f, z = 0.f, tr = 0.f;
            // a stencil spans at most 2 x 2 grid tiles: their list positions
            // (pool) are read once (a tile the stencil does not reach reads
            // its base tile's entry)
            int gp[2][2] = {{0, 0}, {0, 0}};
            if (POOL) {
                const int t0 = sl * geo.ntd + (ib >> 4) * geo.nty + (jb >> 4);
                const int ox1 = ((ib + 2) >> 4) != (ib >> 4) ? geo.nty : 0;
                const int oy1 = ((jb + 2) >> 4) != (jb >> 4) ? 1 : 0;
                gp[0][0] = gtile_pos[t0];
                gp[1][0] = gtile_pos[t0 + ox1];
                gp[0][1] = gtile_pos[t0 + oy1];
                gp[1][1] = gtile_pos[t0 + ox1 + oy1];
            }
            // The weights factor by axis, w = wx[di] wy[dj]: each stencil row is
            // first summed with wy (and dwy for the strain trace), then the
            // rows with wx (dwx):  tr = sum dwx wy vx + wx dwy vy.
#pragma unroll
            for (int di = 0; di < 3; ++di) {
                // the stencil row's first node; the row's 3 nodes follow it
                const IX k0 = (IX)(ib + di) * (IX)ny + (IX)jb;
                const float4* vfr = POOL ? vf : &el<O32>(vf, (IX)sl * nn + k0);
                float rx = 0.f, ry = 0.f, rz = 0.f, rw = 0.f, ryd = 0.f;
#pragma unroll
                for (int dj = 0; dj < 3; ++dj) {
                    float4 q;
                    if (POOL) {
                        const int i = ib + di, j = jb + dj;
                        const int qx = (i >> 4) - (ib >> 4), qy = (j >> 4) - (jb >> 4);
                        q = vf[(size_t)gp[qx][qy] * TILE_CELLS + (((i & 15) << 4) | (j & 15))];
                    } else {
                        q = vfr[dj];
                    }
                    rx += wy[dj] * q.x;
                    ry += wy[dj] * q.y;
                    rz += wy[dj] * q.z;
                    rw += wy[dj] * q.w;
                    ryd += dwy[dj] * q.y;
                }
                picx += wx[di] * rx;
                picy += wx[di] * ry;
                flx += wx[di] * rz;
                fly += wx[di] * rw;
                tr += dwx[di] * rx + wx[di] * ryd;
            }
            // the elevation, on output steps, after the velocity gathers (its
            // loads do not add to their registers)
            const bool zl = (POOL ? slots[sl].out_step : zo0) != 0;
            if (zl) {
                // one row at a time: output steps are rare, and the rows' loads
                // in flight together would add registers to every step
#pragma unroll 1
                for (int di = 0; di < 3; ++di) {
                    const float* tzr = &el<O32>(tz, (IX)(ib + di) * (IX)ny + (IX)jb);
                    float rt = 0.f;
#pragma unroll
                    for (int dj = 0; dj < 3; ++dj) rt += wy[dj] * tzr[dj];
                    z += wx[di] * rt;
                }
            }
            tr *= dt;
            const float nvx = fr * (vxp + flx) + (1.f - fr) * picx;
            const float nvy = fr * (vyp + fly) + (1.f - fr) * picy;
            float nox = ox + dt * picx * inv, noy = oy + dt * picy * inv;
            const float fxs = floorf(nox), fys = floorf(noy);
            int nci = ci + (int)fxs, ncj = cj + (int)fys;
            nox -= fxs;
            noy -= fys;
            if (nox >= 1.f) { nox -= 1.f; nci += 1; }
            if (noy >= 1.f) { noy -= 1.f; ncj += 1; }
            if (!(nox >= 0.f)) nox = 0.f; // NaN guard: keeps indices in range
            if (!(noy >= 0.f)) noy = 0.f;
            if (nci < 0) { nci = 0; nox = 0.f; }
            if (nci > nx - 1) { nci = nx - 1; nox = 0.f; }
            if (ncj < 0) { ncj = 0; noy = 0.f; }
            if (ncj > ny - 1) { ncj = ny - 1; noy = 0.f; }
            cell = (nci << 16) | ncj;
            ox = nox;
            oy = noy;
            vxp = nvx;
            vyp = nvy;
            if (vzl) vzp = nvx * Gx + nvy * Gy;
            vzset = vzl;
            hp = hp / (1.f + tr);
            zp = z;
            zset = zl;
            if (cell_out(geo, nci, ncj)) { // outflow: frozen from here on
                out = true;
                parked = true;
                vxp = vyp = vzp = 0.f;
                vzset = true;
            }
        } else if (POOL) {
            vzp = el<O32>(P.vz, s); // a batch rewrites vz every step
            vzset = true;
        }
        // vz and the elevation are carried (re-sort steps move them) unless
        // this step set them
        if (!vzset && reo) vzp = el<O32>(P.vz, s);
        if (!zset && reo) zp = el<O32>(P.z, s);
        el<O32>(Q.cell, o) = cell;
        el<O32>(Q.o, o) = make_float2(ox, oy);
        el<O32>(Q.v, o) = make_float2(vxp, vyp);
        if (vzset || reo) el<O32>(Q.vz, o) = vzp;
        el<O32>(Q.h, o) = hp;
        if (zset || reo) el<O32>(Q.z, o) = zp;
        if (reo) {
            el<O32>(Q.m, o) = el<O32>(P.m, s);
            el<O32>(Q.pid, o) = el<O32>(P.pid, s);
        }
        if (reo || parked) el<O32>(Q.slot, o) = sl | (out ? SLOT_OUT : 0);
        kk = ekey(geo, st, sl, out, cell, ox, oy, &t, &cb);
        pub = slot_active(st) && !out;
        v2 = vxp * vxp + vyp * vyp;
    }
    wave_count_any(t, in && kk >= 0, tile_count);
    if (in) {
        el<O32>(key, o) = kk;
        el<O32>(kcell, o) = (uint8_t)cb;
    }
    publish_vmax(sl, pub, v2, vmx);
}

// --- downloads / diagnostics -------------------------------------------------

// Compacts one slot's particles of the current buffer into out[0..] (order
// arbitrary; the host scatters by creation index).
__global__ void k_ecollect(ECtl* c, PB32 A, PB32 B, int slot, int* __restrict__ counter, int32_t* __restrict__ o_cell,
                           float* __restrict__ o_f, int32_t* __restrict__ o_pid, int cap) {
    const PB32 P = ecur(c, A, B);
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= c->n_store) return;
    if ((P.slot[s] & SLOT_MASK) != slot) return;
    const int q = atomicAdd(counter, 1);
    if (q >= cap) return;
    o_cell[q] = P.cell[s];
    o_pid[q] = P.pid[s];
    float* f = o_f + (size_t)q * 8;
    const float2 po = P.o[s], pv = P.v[s];
    f[0] = po.x;
    f[1] = po.y;
    f[2] = pv.x;
    f[3] = pv.y;
    f[4] = P.vz[s];
    f[5] = P.z[s];
    f[6] = P.h[s];
    f[7] = P.m[s];
}

// --- the download that allocates nothing --------------------------------------

// Fixed scratch of slab_batch_download_particles_noalloc: a counter, then the
// nine FP64 fields of at most NOALLOC_MAX particles in creation order.  Each
// GPU the library runs on holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (scratch_address), never through the name.
constexpr int NOALLOC_MAX = SLAB_NOALLOC_MAX_PARTICLES;
constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_HEAD + 9 * (size_t)NOALLOC_MAX * sizeof(double);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

__global__ void k_noalloc_clear(int* __restrict__ counter) {
    if (threadIdx.x == 0) *counter = 0;
}

// One slot's particles of the current buffer as the FP64 fields of
// slab_batch_download_particles, written at their creation index:
// out[f * NOALLOC_MAX + pid] for f = x, y, vx, vy, vz, z, h, A, m (the same
// operations as the host side of engine_particles, so the same values).
__global__ void __launch_bounds__(WG) k_ecollect_pid(const ECtl* c, PB32 A, PB32 B, int slot, int n, double rho,
                                                     int* __restrict__ counter, double* __restrict__ out) {
    const PB32 P = ecur(c, A, B);
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= c->n_store) return;
    if ((P.slot[s] & SLOT_MASK) != slot) return;
    const int q = P.pid[s];
    if (q < 0 || q >= n) return;
    atomicAdd(counter, 1);
    const int cell = P.cell[s];
    const float2 po = P.o[s], pv = P.v[s];
    const float hp = P.h[s], mp = P.m[s];
    const double dx = c->dx;
    out[q] = ((double)(cell >> 16) + (double)po.x) * dx;
    out[(size_t)NOALLOC_MAX + q] = ((double)(cell & 0xffff) + (double)po.y) * dx;
    out[2 * (size_t)NOALLOC_MAX + q] = pv.x;
    out[3 * (size_t)NOALLOC_MAX + q] = pv.y;
    out[4 * (size_t)NOALLOC_MAX + q] = P.vz[s];
    out[5 * (size_t)NOALLOC_MAX + q] = P.z[s];
    out[6 * (size_t)NOALLOC_MAX + q] = hp;
    out[7 * (size_t)NOALLOC_MAX + q] = (double)mp / ((double)hp * rho);
    out[8 * (size_t)NOALLOC_MAX + q] = mp;
}

// Bounding box of one slot's particles in the buffer its last P2G read
// (for the reference's remesh window, src/simulation.cpp:594-606).
__global__ void k_ebbox(ECtl* c, PB32 A, PB32 B, int slot, int n, unsigned long long* __restrict__ bb) {
    const PB32 P = c->p2g_parity ? B : A;
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= n || (P.slot[s] & SLOT_MASK) != slot) return;
    const int cell = P.cell[s];
    const float2 po = P.o[s];
    const double x = ((double)(cell >> 16) + (double)po.x) * c->dx;
    const double y = ((double)(cell & 0xffff) + (double)po.y) * c->dx;
    atomicMin(&bb[0], ord_bits(x));
    atomicMax(&bb[1], ord_bits(x));
    atomicMin(&bb[2], ord_bits(y));
    atomicMax(&bb[3], ord_bits(y));
}

__global__ void k_ewindow(const ECtl* c, int slot, int i0, int j0, int wnx, int wny, const int* __restrict__ gstamp,
                          const float4* __restrict__ vf, const float* __restrict__ gm, const float* __restrict__ gh,
                          const float* __restrict__ gvz, double* __restrict__ o) {
    const int q = blockIdx.x * WG + threadIdx.x;
    const size_t wn = (size_t)wnx * wny;
    if ((size_t)q >= wn) return;
    const int a = q / wny, b = q - a * wny;
    const int i = i0 + a, j = j0 + b;
    const int t = slot * c->ntd + (i >> 4) * c->nty + (j >> 4);
    double h = 0, vx = 0, vy = 0, vz = 0, m = 0;
    if (gstamp[t] == c->epoch) {
        const size_t k = (size_t)slot * c->nx * c->ny + (size_t)i * c->ny + j;
        const float4 v = vf[k];
        vx = v.x;
        vy = v.y;
        if (gm) {
            h = gh[k];
            vz = gvz[k];
            m = gm[k];
        }
    }
    o[q] = h;
    o[wn + q] = vx;
    o[2 * wn + q] = vy;
    o[3 * wn + q] = vz;
    o[4 * wn + q] = m;
}

// FP64 sums over one slot's current grid tiles and particles.
__global__ void __launch_bounds__(WG) k_ediag(const ECtl* c, PB32 A, PB32 B, int slot, const int* __restrict__ gtiles,
                                              const float4* __restrict__ vf, const float* __restrict__ gm,
                                              double rho, double* __restrict__ part) {
    __shared__ double sh[6][WG];
    double a[6] = {0, 0, 0, 0, 0, 0};
    const int nty = c->nty, ntd = c->ntd;
    const size_t nn = (size_t)c->nx * c->ny;
    if (gm) {
        for (int w = blockIdx.x; w < c->n_gtiles; w += gridDim.x) {
            const int t = gtiles[w];
            if (t / ntd != slot) continue;
            const int local = t - slot * ntd;
            const int tx = local / nty, ty = local - tx * nty;
            const int i = tx * TILE + (threadIdx.x >> 4), j = ty * TILE + (threadIdx.x & 15);
            if (i < c->nx && j < c->ny) {
                const size_t k = (size_t)slot * nn + (size_t)i * c->ny + j;
                const double m = gm[k];
                const float4 v = vf[k];
                a[0] += m;
                if (m > 1e-6) {
                    a[1] += m * (double)v.x;
                    a[2] += m * (double)v.y;
                    a[3] += 1.0;
                }
            }
        }
    }
    const PB32 P = ecur(c, A, B);
    for (int r = blockIdx.x * WG + threadIdx.x; r < c->n_store; r += gridDim.x * WG) {
        if ((P.slot[r] & SLOT_MASK) != slot) continue;
        const double m = P.m[r];
        a[4] += m;
        a[5] += m / rho;
    }
    for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int o = WG / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) part[6 * blockIdx.x + q] = sh[q][0];
}

// Returns a released slot's max-map pool tiles, zeroed, to the free stack
// (one workgroup per terrain tile of the slot; no allocation runs meanwhile).
__global__ void __launch_bounds__(WG) k_hm_release(ECtl* c, int slot, int* __restrict__ hm_map,
                                                   int* __restrict__ hm_stack, float* __restrict__ h_max,
                                                   float* __restrict__ v_max2) {
    for (int local = blockIdx.x; local < c->ntd; local += gridDim.x) {
        const int t = slot * c->ntd + local;
        const int idx = hm_map[t];
        if (idx < 0) continue;
        h_max[(size_t)idx * TILE_CELLS + threadIdx.x] = 0.f;
        v_max2[(size_t)idx * TILE_CELLS + threadIdx.x] = 0.f;
        if (threadIdx.x == 0) {
            const int k = atomicSub(&c->hm_top, 1) - 1;
            hm_stack[k] = idx;
            hm_map[t] = -1;
        }
    }
}

// One slot's max maps from the pool as dense terrain arrays (0 where the slot
// never reached).
__global__ void k_hm_dense(const ECtl* c, int slot, const int* __restrict__ hm_map, const float* __restrict__ h_max,
                           const float* __restrict__ v_max2, float* __restrict__ oh, float* __restrict__ ov) {
    const size_t nn = (size_t)c->nx * c->ny;
    const size_t q = (size_t)blockIdx.x * WG + threadIdx.x;
    if (q >= nn) return;
    const int i = (int)(q / c->ny), j = (int)(q - (size_t)i * c->ny);
    const int idx = hm_map[slot * c->ntd + (i >> 4) * c->nty + (j >> 4)];
    float h = 0.f, v = 0.f;
    if (idx >= 0) {
        const size_t k = (size_t)idx * TILE_CELLS + (((i & 15) << 4) | (j & 15));
        h = h_max[k];
        v = v_max2[k];
    }
    oh[q] = h;
    ov[q] = v;
}

// ============================================================== host code

namespace {

constexpr int kPersistE = 2048;
constexpr int kReorderEvery = 16; // default storage re-sort period (steps)
constexpr int kNumK = 10;

bool hip_ok(hipError_t e, const char* what, std::string* err) {
    if (e == hipSuccess) return true;
    if (err) *err = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return false;
}

} // namespace

struct Engine32 {
    hipStream_t stream = nullptr;
    ECtl* d_c = nullptr;
    ECtl* h_c = nullptr;
    Slot* d_slots = nullptr;
    Slot* h_slots = nullptr;
    int nslots = 1, nx = 0, ny = 0, ntx = 0, nty = 0, ntd = 0, ntiles = 0, nblk = 0;
    // up to this many scan blocks, k_eoffsets forms the block offsets itself
    // (each block reads all block sums) instead of a k_escan launch
    int scan_merge_max = 2048;
    size_t nn = 0;
    double dx = 0, inv = 0;
    bool single = true;
    int64_t cap = 0;
    int tile_cap = 0;
    float *tz = nullptr, *tgx = nullptr, *tgy = nullptr, *tcth = nullptr;
    float *hmax = nullptr, *vmax2 = nullptr;
    float4* vf = nullptr;
    float *om = nullptr, *oh = nullptr, *ovz = nullptr;
    unsigned long long* vmx = nullptr;
    int *tile_count = nullptr, *cursor = nullptr, *tile_begin = nullptr, *pstamp = nullptr, *gstamp = nullptr,
        *tile_pos = nullptr;
    int *ptiles = nullptr, *gtiles = nullptr, *bsum = nullptr, *boff = nullptr;
    int2* pinfo = nullptr; // per particle-tile list position: {begin, count}
    int* cell_begin = nullptr;
    uint8_t* tdefer = nullptr; // per tile list position: left by the small-tile pass of k_esp
    float* tpart = nullptr;
    PB32 A{}, B{};
    int *key = nullptr, *idx = nullptr, *tmp = nullptr;
    uint8_t* kcell = nullptr; // per stored particle: its cell in its tile (the key holds the tile)
    // stored particles the particle-range grids (G2P, scatter) are sized for:
    // the live count rounded up to a power of two, so a lightly filled
    // capacity launches few idle blocks (a captured step is rebuilt when it
    // no longer fits)
    int64_t launch_n = 0;
    int grid_wg = 8192; // workgroups of the grid kernel (sized from the grid-tile count)
    double* dpart = nullptr;
    std::vector<void*> allocs, pallocs;
    std::vector<int64_t> admit_base; // storage offset of each slot's particles at admission
    // timing / graph
    bool timing = false;
    struct Pend {
        int kid, step;
        hipEvent_t a, b;
    };
    std::vector<Pend> pending;
    std::vector<hipEvent_t> ev_free;
    double ms[kNumK] = {};
    int64_t cnt[kNumK] = {};
    bool use_graph = false;
    int persist = kPersistE; // workgroups of the tile sort + P2G kernel
    int sort_cap = ES_CAPS[2]; // in-LDS capacity of the tile sort (particles)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int device = 0;
    unsigned char* scratch_ptr = nullptr; // this device's address of `scratch`
    // batches keep the grid velocity in a pool of gtile_cap grid tiles
    // (single runs: dense over the terrain)
    bool pool_vf = false;
    // a single run whose node and particle arrays all stay below 4 GiB: G2P
    // addresses them with 32-bit byte offsets
    bool off32 = false;
    int gtile_cap = 0;
    int* gtile_pos = nullptr;
    // admission assumes a scenario's active tiles may grow to this many
    // times those of its release
    int tile_growth = 4;
    std::vector<int64_t> res_ptiles, res_gtiles; // per slot: tiles reserved at admission
    // batches keep the max maps in a pool of touched tiles (single runs: dense)
    int hm_cap = 0;
    int *hm_map = nullptr, *hm_stack = nullptr;
};

namespace {

template <typename T>
bool ealloc(std::vector<void*>& list, T** p, size_t n, std::string* err) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        if (err) *err = "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed";
        return false;
    }
    list.push_back(q);
    *p = (T*)q;
    return true;
}

void efree(std::vector<void*>& l) {
    for (void* p : l) (void)hipFree(p);
    l.clear();
}

bool push_all(Engine32* e, std::string* err) {
    if (!hip_ok(hipMemcpyAsync(e->d_c, e->h_c, sizeof(ECtl), hipMemcpyHostToDevice, e->stream), "ctl upload", err))
        return false;
    if (!hip_ok(hipMemcpyAsync(e->d_slots, e->h_slots, sizeof(Slot) * e->nslots, hipMemcpyHostToDevice, e->stream),
                "slot upload", err))
        return false;
    return hip_ok(hipStreamSynchronize(e->stream), "sync", err);
}

hipEvent_t ev_get(Engine32* e) {
    if (!e->ev_free.empty()) {
        hipEvent_t x = e->ev_free.back();
        e->ev_free.pop_back();
        return x;
    }
    hipEvent_t x = nullptr;
    (void)hipEventCreate(&x);
    return x;
}

template <typename F>
void tk(Engine32* e, int kid, int step, F&& f) {
    if (!e->timing) {
        f();
        return;
    }
    hipEvent_t a = ev_get(e), b = ev_get(e);
    (void)hipEventRecord(a, e->stream);
    f();
    (void)hipEventRecord(b, e->stream);
    e->pending.push_back({kid, step, a, b});
}

// at least n blocks, a multiple of N_XCD (for xcd_block)
unsigned xcd_grid(int64_t n) { return (unsigned)((std::max<int64_t>(1, n) + N_XCD - 1) / N_XCD * N_XCD); }

void enqueue_estep(Engine32* e, int step) {
    hipStream_t s = e->stream;
    ECtl* c = e->d_c;
    const unsigned tb = (unsigned)std::max(1, std::min(e->ntiles, e->grid_wg));
    const unsigned pb = (unsigned)std::max(1, std::min(e->tile_cap, e->persist));
    const int many = e->nslots > 4;
    if (many) tk(e, 0, step, [&] { k_evmax<<<e->nslots, 64, 0, s>>>(c, e->vmx); });
    tk(e, 1, step, [&] {
        k_ereduce<<<e->nblk + 1, WG, 0, s>>>(c, e->d_slots, e->tile_count, e->ntiles, e->bsum, e->vmx, many);
    });
    const int merged = e->nblk <= e->scan_merge_max;
    if (!merged) tk(e, 2, step, [&] { k_escan<<<1, WG, 0, s>>>(c, e->bsum, e->nblk, e->boff); });
    tk(e, 3, step, [&] {
        k_eoffsets<<<e->nblk, WG, 0, s>>>(c, e->d_slots, e->tile_count, e->ntiles, e->boff, e->bsum, e->nblk, merged,
                                           e->tile_begin, e->ptiles, e->gtiles, e->pstamp, e->gstamp, e->tile_pos,
                                           e->gtile_pos, e->hm_map, e->hm_stack, e->pinfo);
    });
    tk(e, 4, step, [&] {
        const unsigned nb = (unsigned)std::max<int64_t>(1, (e->launch_n + WG * SC_PER - 1) / (WG * SC_PER));
        k_escatter<<<nb, WG, 0, s>>>(c, e->key, e->tile_begin, e->cursor, e->tmp);
    });
    // the tile sort and P2G run as one launch (k_esp), timed as "p2g"
    tk(e, 6, step, [&] {
        if (e->sort_cap == ES_CAPS[0])
            k_esp<ES_CAPS[0]><<<pb, WG, 0, s>>>(c, e->d_slots, e->A, e->B, e->kcell, e->ptiles, e->pinfo, e->tile_count,
                                                e->cursor, e->tmp, e->idx, e->cell_begin, e->tdefer, e->tgx, e->tgy,
                                                e->tcth, e->tpart);
        else if (e->sort_cap == ES_CAPS[1])
            k_esp<ES_CAPS[1]><<<pb, WG, 0, s>>>(c, e->d_slots, e->A, e->B, e->kcell, e->ptiles, e->pinfo, e->tile_count,
                                                e->cursor, e->tmp, e->idx, e->cell_begin, e->tdefer, e->tgx, e->tgy,
                                                e->tcth, e->tpart);
        else
            k_esp<ES_CAPS[2]><<<pb, WG, 0, s>>>(c, e->d_slots, e->A, e->B, e->kcell, e->ptiles, e->pinfo, e->tile_count,
                                                e->cursor, e->tmp, e->idx, e->cell_begin, e->tdefer, e->tgx, e->tgy,
                                                e->tcth, e->tpart);
    });
    tk(e, 7, step, [&] {
        if (e->pool_vf)
            k_egrid<true><<<tb, WG, 0, s>>>(c, e->d_slots, e->gtiles, e->pstamp, e->tile_pos, e->tpart, e->tgx, e->tgy,
                                            e->tcth, e->hmax, e->vmax2, e->vf, e->om, e->oh, e->ovz, e->hm_map);
        else
            k_egrid<false><<<tb, WG, 0, s>>>(c, e->d_slots, e->gtiles, e->pstamp, e->tile_pos, e->tpart, e->tgx, e->tgy,
                                             e->tcth, e->hmax, e->vmax2, e->vf, e->om, e->oh, e->ovz, nullptr);
    });
    tk(e, 8, step, [&] {
        const unsigned gb = xcd_grid((e->launch_n + WG - 1) / WG);
        if (e->pool_vf)
            k_eg2p<true, false><<<gb, WG, 0, s>>>(c, e->d_slots, e->A, e->B, e->idx, e->vf, e->gtile_pos, e->tz,
                                                  e->tgx, e->tgy, e->key, e->kcell, e->tile_count, e->vmx);
        else if (e->off32)
            k_eg2p<false, true><<<gb, WG, 0, s>>>(c, e->d_slots, e->A, e->B, e->idx, e->vf, e->gtile_pos, e->tz,
                                                  e->tgx, e->tgy, e->key, e->kcell, e->tile_count, e->vmx);
        else
            k_eg2p<false, false><<<gb, WG, 0, s>>>(c, e->d_slots, e->A, e->B, e->idx, e->vf, e->gtile_pos, e->tz,
                                                   e->tgx, e->tgy, e->key, e->kcell, e->tile_count, e->vmx);
    });
}

// Ends the last enqueued step (a no-op when none is pending); enqueued before
// the host reads the control block.
void enqueue_eend(Engine32* e, int step) {
    tk(e, 9, step, [&] { k_eend<<<1, WG, 0, e->stream>>>(e->d_c, e->d_slots); });
}

void drop_graph(Engine32* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
}

// The grid kernel's workgroups: about one per two grid tiles of the last step,
// a power of two in [8192, 32768] (measured on MI355X: the 8M-particle run's
// ~8.5 k grid tiles 45.0 / 41.8 / 42.1 / 46.6 us with 2,048 / 8,192 / 16,384 /
// 32,768 workgroups; the 256-scenario batch's ~70 k tiles 223 / 199 / 194 /
// 191 us).  It grows when the tiles outgrow it and shrinks below a quarter.
void size_grid_workgroups(Engine32* e) {
    const int64_t half = std::max<int64_t>(1, e->h_c->n_gtiles / 2);
    if (half <= e->grid_wg && half * 4 >= e->grid_wg) return;
    int want = 8192;
    while (want < half && want < 32768) want <<= 1;
    if (want == e->grid_wg) return;
    e->grid_wg = want;
    drop_graph(e);
}

// Re-sizes the particle-range grids to the stored count when it outgrew them
// or fell below a quarter (the captured step is dropped and rebuilt).
void size_particle_grids(Engine32* e) {
    size_grid_workgroups(e);
    const int64_t need = std::max<int64_t>(1, e->h_c->n_store);
    if (need <= e->launch_n && need * 4 >= e->launch_n) return;
    int64_t want = 65536;
    while (want < need) want <<= 1;
    e->launch_n = std::min<int64_t>(e->cap, want);
    if (e->launch_n < need) e->launch_n = need;
    drop_graph(e);
}

bool build_graph(Engine32* e, std::string* err) {
    drop_graph(e);
    if (!hip_ok(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "begin capture", err)) return false;
    const bool tm = e->timing;
    e->timing = false;
    enqueue_estep(e, 0);
    e->timing = tm;
    if (!hip_ok(hipStreamEndCapture(e->stream, &e->graph), "end capture", err)) return false;
    return hip_ok(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0), "graph instantiate", err);
}

void collect(Engine32* e, int valid_steps) {
    for (auto& p : e->pending) {
        if (p.step < valid_steps) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                e->ms[p.kid] += ms;
                e->cnt[p.kid] += 1;
            }
        }
        e->ev_free.push_back(p.a);
        e->ev_free.push_back(p.b);
    }
    e->pending.clear();
}

// terrainGradients + terrainTheta (src/complex_terrain.cpp:84-111), host FP64
// with the host libm, stored as FP32 tables {Gx, Gy, cos(theta)}.
void terrain32(const double* z, int nx, int ny, double inv, std::vector<float>& gx, std::vector<float>& gy,
               std::vector<float>& cth) {
    const size_t nn = (size_t)nx * ny;
    gx.assign(nn, 0.f);
    gy.assign(nn, 0.f);
    cth.assign(nn, 1.f);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            for (int j = 0; j < ny; ++j) {
                const size_t k = (size_t)i * ny + j;
                double dzx = 0.0, dzy = 0.0;
                if (i < nx - 1 && j < ny - 1) {
                    const double z00 = z[k], z01 = z[k + 1], z10 = z[k + ny], z11 = z[k + ny + 1];
                    dzx = (z10 + z11 - z00 - z01) / 2. * inv;
                    dzy = (z01 + z11 - z00 - z10) / 2. * inv;
                }
                gx[k] = (float)dzx;
                gy[k] = (float)dzy;
                cth[k] = (float)std::cos(std::atan(std::sqrt(dzx * dzx + dzy * dzy)));
            }
        }
    };
    std::vector<std::thread> ts;
    int per = (nx + (int)nth - 1) / (int)nth;
    for (unsigned t = 0; t < nth; ++t) {
        int a = (int)t * per, b = std::min(nx, a + per);
        if (a < b) ts.emplace_back(work, a, b);
    }
    for (auto& t : ts) t.join();
}

} // namespace

int64_t engine_particle_bytes() {
    // two state buffers (cell, ox, oy, vx, vy, vz, h, m, z, pid, slot = 44 B) + key, cell byte, idx, tmp
    return 2 * 44 + 13;
}

int64_t engine_slot_bytes(const Engine32* e) {
    // max maps (h_max, v_max2) per node; a single run also keeps its grid
    // velocity and output fields dense
    if (e->single) return (int64_t)e->nn * (8 + 16 + 12) + VMAX_SHARDS * 8 + (int64_t)sizeof(Slot);
    // a batch slot: its control block, dt shards and tile -> max-map pool index
    return (int64_t)e->ntd * 4 + VMAX_SHARDS * 8 + (int64_t)sizeof(Slot);
}

bool engine_set_tile_growth(Engine32* e, int factor, std::string* err) {
    if (factor < 0 || factor > 1024) {
        if (err) *err = "tile growth factor must be in [0, 1024]";
        return false;
    }
    e->tile_growth = factor;
    return true;
}

void engine_tiles(const Engine32* e, int64_t out[9]) {
    out[0] = e->gtile_cap;
    out[1] = e->h_c->n_gtiles;
    out[2] = e->tile_cap;
    out[3] = e->h_c->n_ptiles;
    out[4] = e->pool_vf ? (int64_t)TILE_CELLS * (int64_t)sizeof(float4) : 0;
    out[5] = (int64_t)CELL_STRIDE * 4 + (int64_t)P2G_FIELDS * REGION_NODES * 4;
    out[6] = e->hm_cap;
    out[7] = e->single ? 0 : e->h_c->hm_top;
    out[8] = e->single ? 0 : (int64_t)TILE_CELLS * 8;
}

// Device address of `scratch` on each device, taken once per device with that
// device current.
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[64] = {};

bool scratch_address(int dev, unsigned char** out, std::string* err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= 64) {
        if (err) *err = "device id out of range for the download scratch";
        return false;
    }
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !p) {
            (void)hipGetLastError();
            if (err) *err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(r);
            return false;
        }
        g_scratch_addr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_addr[dev];
    return true;
}

Engine32* engine_create(int nx, int ny, double dx, const double* z, int nslots, int64_t max_particles, bool single,
                        std::string* err) {
    if (nx < 4 || ny < 4 || nx > 65535 || ny > 65535) {
        if (err) *err = "FP32 engine: terrain must be between 4 and 65535 nodes per side";
        return nullptr;
    }
    if (nslots < 1 || max_particles < 1 || max_particles > ((int64_t)1 << 30)) {
        if (err) *err = "FP32 engine: need >= 1 slot and 1..2^30 particles of capacity";
        return nullptr;
    }
    Engine32* e = new Engine32();
    e->nslots = nslots;
    e->single = single;
    e->nx = nx;
    e->ny = ny;
    e->nn = (size_t)nx * ny;
    e->dx = dx;
    e->inv = 1. / dx;
    e->ntx = (nx + TILE - 1) / TILE;
    e->nty = (ny + TILE - 1) / TILE;
    e->ntd = e->ntx * e->nty;
    e->ntiles = nslots * (e->ntd + 1);
    if ((double)e->ntiles >= 2147483647.0) {
        if (err) *err = "FP32 engine: slots x terrain tiles exceed the 31-bit tile ids";
        delete e;
        return nullptr;
    }
    e->nblk = (e->ntiles + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (const char* v = std::getenv("SLAB_SCAN_MERGE_MAX_BLOCKS")) e->scan_merge_max = std::atoi(v);
    e->cap = max_particles;
    e->tile_cap = (int)std::min<int64_t>(e->ntiles, std::max<int64_t>(65536, max_particles / 32));
    // a batch's grid tiles are its particle tiles plus right / upper /
    // diagonal neighbours: at most 4x, in practice well under 2x
    e->pool_vf = !single;
    e->off32 = single && (int64_t)e->nn * (int64_t)sizeof(float4) < ((int64_t)1 << 32) &&
               max_particles * (int64_t)sizeof(float4) < ((int64_t)1 << 32);
    // the 64-bit addressing of larger runs, on any run (tests of that path)
    if (const char* v = std::getenv("SLAB_G2P_OFFSET64"))
        if (std::atoi(v) != 0) e->off32 = false;
    e->gtile_cap = single ? e->ntiles : (int)std::min<int64_t>(e->ntiles, 2 * (int64_t)e->tile_cap);
    // smaller active-tile capacities (tests of admission and the overflow path)
    if (const char* v = std::getenv("SLAB_TILE_CAP")) {
        e->tile_cap = std::max(1, std::min(e->tile_cap, std::atoi(v)));
        if (!single) e->gtile_cap = std::max(1, std::min(e->gtile_cap, 2 * e->tile_cap));
    }
    // a batch's max maps: a pool of touched tiles, sized like the grid-tile
    // pool but at least 2^20 tiles (2 GB), at most every terrain tile of every slot
    if (!single) {
        const int64_t live = (int64_t)nslots * e->ntd;
        e->hm_cap = (int)std::min<int64_t>(live, std::max<int64_t>(e->gtile_cap, (int64_t)1 << 20));
        if (const char* v = std::getenv("SLAB_MAXMAP_TILE_CAP")) e->hm_cap = std::max(1, std::min(e->hm_cap, std::atoi(v)));
    }
    std::string why;
    auto fail = [&](const std::string& m) -> Engine32* {
        if (err) *err = m;
        engine_destroy(e);
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return fail("hipStreamCreate failed");
    if (hipGetDevice(&e->device) != hipSuccess) return fail("hipGetDevice failed");
    if (!scratch_address(e->device, &e->scratch_ptr, &why)) return fail(why);
    if (hipHostMalloc((void**)&e->h_c, sizeof(ECtl), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&e->h_slots, sizeof(Slot) * nslots, hipHostMallocDefault) != hipSuccess)
        return fail("hipHostMalloc failed");
    std::memset(e->h_c, 0, sizeof(ECtl));
    std::memset(e->h_slots, 0, sizeof(Slot) * nslots);
    const size_t sn = (size_t)nslots * e->nn;
    bool ok = ealloc(e->allocs, &e->d_c, 1, &why) && ealloc(e->allocs, &e->d_slots, nslots, &why) &&
              ealloc(e->allocs, &e->tz, e->nn, &why) && ealloc(e->allocs, &e->tgx, e->nn, &why) &&
              ealloc(e->allocs, &e->tgy, e->nn, &why) && ealloc(e->allocs, &e->tcth, e->nn, &why) &&
              ealloc(e->allocs, &e->hmax, single ? sn : (size_t)e->hm_cap * TILE_CELLS, &why) &&
              ealloc(e->allocs, &e->vmax2, single ? sn : (size_t)e->hm_cap * TILE_CELLS, &why) &&
              ealloc(e->allocs, &e->vf, e->pool_vf ? (size_t)e->gtile_cap * TILE_CELLS : sn, &why) &&
              ealloc(e->allocs, &e->gtile_pos, e->ntiles, &why) &&
              ealloc(e->allocs, &e->vmx, (size_t)nslots * VMAX_SHARDS, &why) &&
              ealloc(e->allocs, &e->tile_count, e->ntiles, &why) && ealloc(e->allocs, &e->cursor, e->ntiles, &why) &&
              ealloc(e->allocs, &e->tile_begin, e->ntiles, &why) &&
              ealloc(e->allocs, &e->pstamp, e->ntiles, &why) && ealloc(e->allocs, &e->gstamp, e->ntiles, &why) &&
              ealloc(e->allocs, &e->tile_pos, e->ntiles, &why) && ealloc(e->allocs, &e->ptiles, e->ntiles, &why) &&
              ealloc(e->allocs, &e->pinfo, e->ntiles, &why) &&
              ealloc(e->allocs, &e->gtiles, e->ntiles, &why) && ealloc(e->allocs, &e->bsum, 3 * (size_t)e->nblk, &why) &&
              ealloc(e->allocs, &e->boff, 3 * (size_t)e->nblk, &why) &&
              ealloc(e->allocs, &e->cell_begin, (size_t)e->tile_cap * CELL_STRIDE, &why) &&
              ealloc(e->allocs, &e->tdefer, (size_t)e->tile_cap, &why) &&
              ealloc(e->allocs, &e->tpart, (size_t)e->tile_cap * P2G_FIELDS * REGION_NODES, &why) &&
              ealloc(e->allocs, &e->dpart, 6 * 256, &why);
    if (ok && single)
        ok = ealloc(e->allocs, &e->om, sn, &why) && ealloc(e->allocs, &e->oh, sn, &why) &&
             ealloc(e->allocs, &e->ovz, sn, &why);
    for (PB32* b : {&e->A, &e->B}) {
        ok = ok && ealloc(e->pallocs, &b->cell, max_particles, &why);
        ok = ok && ealloc(e->pallocs, &b->o, max_particles, &why) && ealloc(e->pallocs, &b->v, max_particles, &why);
        for (float** q : {&b->vz, &b->h, &b->m, &b->z})
            ok = ok && ealloc(e->pallocs, q, max_particles, &why);
        ok = ok && ealloc(e->pallocs, &b->pid, max_particles, &why) && ealloc(e->pallocs, &b->slot, max_particles, &why);
    }
    ok = ok && ealloc(e->pallocs, &e->key, max_particles, &why) && ealloc(e->pallocs, &e->kcell, max_particles, &why) &&
         ealloc(e->pallocs, &e->idx, max_particles, &why) &&
         ealloc(e->pallocs, &e->tmp, max_particles, &why);
    if (ok && !single)
        ok = ealloc(e->allocs, &e->hm_map, (size_t)nslots * e->ntd, &why) &&
             ealloc(e->allocs, &e->hm_stack, (size_t)e->hm_cap, &why);
    if (!ok) return fail(why);
    if (!single) {
        std::vector<int32_t> st(e->hm_cap);
        for (int q = 0; q < e->hm_cap; ++q) st[q] = q;
        ok = hipMemsetAsync(e->hm_map, 0xff, (size_t)nslots * e->ntd * 4, e->stream) == hipSuccess &&
             hipMemcpyAsync(e->hm_stack, st.data(), st.size() * 4, hipMemcpyHostToDevice, e->stream) == hipSuccess &&
             hipStreamSynchronize(e->stream) == hipSuccess;
        if (!ok) return fail("max-map pool set-up failed");
    }
    std::vector<float> gx, gy, cth, zf(e->nn);
    terrain32(z, nx, ny, e->inv, gx, gy, cth);
    for (size_t k = 0; k < e->nn; ++k) zf[k] = (float)z[k];
    hipStream_t s = e->stream;
    ok = hipMemcpyAsync(e->tz, zf.data(), e->nn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(e->tgx, gx.data(), e->nn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(e->tgy, gy.data(), e->nn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(e->tcth, cth.data(), e->nn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemsetAsync(e->hmax, 0, (single ? sn : (size_t)e->hm_cap * TILE_CELLS) * 4, s) == hipSuccess &&
         hipMemsetAsync(e->vmax2, 0, (single ? sn : (size_t)e->hm_cap * TILE_CELLS) * 4, s) == hipSuccess &&
         hipMemsetAsync(e->vf, 0, (e->pool_vf ? (size_t)e->gtile_cap * TILE_CELLS : sn) * sizeof(float4), s) ==
             hipSuccess &&
         hipMemsetAsync(e->gtile_pos, 0, (size_t)e->ntiles * 4, s) == hipSuccess &&
         hipMemsetAsync(e->vmx, 0, (size_t)nslots * VMAX_SHARDS * 8, s) == hipSuccess &&
         hipMemsetAsync(e->tile_count, 0, (size_t)e->ntiles * 4, s) == hipSuccess &&
         hipMemsetAsync(e->cursor, 0, (size_t)e->ntiles * 4, s) == hipSuccess &&
         hipMemsetAsync(e->pstamp, 0xff, (size_t)e->ntiles * 4, s) == hipSuccess &&
         hipMemsetAsync(e->gstamp, 0xff, (size_t)e->ntiles * 4, s) == hipSuccess;
    if (ok && single)
        ok = hipMemsetAsync(e->om, 0, sn * 4, s) == hipSuccess && hipMemsetAsync(e->oh, 0, sn * 4, s) == hipSuccess &&
             hipMemsetAsync(e->ovz, 0, sn * 4, s) == hipSuccess;
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) return fail("terrain upload failed");
    ECtl& c = *e->h_c;
    c.nx = nx;
    c.ny = ny;
    c.ntx = e->ntx;
    c.nty = e->nty;
    c.ntd = e->ntd;
    c.nslots = nslots;
    c.single = single ? 1 : 0;
    c.has_out = single ? 1 : 0;
    c.tile_cap = e->tile_cap;
    c.gtile_cap = e->gtile_cap;
    c.hm_cap = e->hm_cap;
    c.hm_top = 0;
    c.reorder_every = kReorderEvery;
    c.dx = dx;
    c.inv_dx = e->inv;
    e->admit_base.assign(nslots, -1);
    e->res_ptiles.assign(nslots, 0);
    e->res_gtiles.assign(nslots, 0);
    if (!push_all(e, err)) {
        engine_destroy(e);
        return nullptr;
    }
    return e;
}

void engine_destroy(Engine32* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    for (auto& p : e->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto x : e->ev_free) (void)hipEventDestroy(x);
    efree(e->pallocs);
    efree(e->allocs);
    if (e->h_c) (void)hipHostFree(e->h_c);
    if (e->h_slots) (void)hipHostFree(e->h_slots);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

bool engine_pull(Engine32* e, std::string* err) {
    if (!hip_ok(hipMemcpyAsync(e->h_c, e->d_c, sizeof(ECtl), hipMemcpyDeviceToHost, e->stream), "ctl download", err))
        return false;
    if (!hip_ok(hipMemcpyAsync(e->h_slots, e->d_slots, sizeof(Slot) * e->nslots, hipMemcpyDeviceToHost, e->stream),
                "slot download", err))
        return false;
    return hip_ok(hipStreamSynchronize(e->stream), "sync", err);
}

const ECtl& engine_ctl(const Engine32* e) { return *e->h_c; }
const Slot& engine_slot(const Engine32* e, int s) { return e->h_slots[s]; }

int64_t engine_free_particles(const Engine32* e) { return e->cap - e->h_c->n_store; }

// A step that stopped on an error (the active tiles exceeded the tile
// capacity) leaves a step half started (slot dt and step counters advanced,
// its end pending, tile lists unwritten): the engine accepts no further step,
// admission or release; downloads still read the last completed step.
bool engine_stopped(const Engine32* e, std::string* err) {
    if (e->h_c->stop != STOP_ERROR) return false;
    if (err) *err = kStoppedByError;
    return true;
}

int engine_admit(Engine32* e, const Slot& prm, int64_t np, const double* x, const double* y, const double* vx,
                 const double* vy, const double* vz, const double* h, const double* A, const double* m,
                 std::string* err) {
    if (!engine_pull(e, err)) return -1;
    if (engine_stopped(e, err)) return -1;
    int slot = -1;
    for (int s = 0; s < e->nslots; ++s)
        if (e->h_slots[s].status == SLOT_FREE) {
            slot = s;
            break;
        }
    if (slot < 0) {
        if (err) *err = "no free slot";
        return -1;
    }
    const int64_t base = e->h_c->n_store;
    if (np < 1 || base + np > e->cap) {
        if (err) *err = "particle capacity exhausted";
        return -1;
    }
    const size_t n = (size_t)np;
    std::vector<int32_t> cell(n), pid(n), sl(n, slot);
    std::vector<float2> fo(n), fv(n);
    std::vector<float> fvz(n), fh(n), fm(n), zz(n, 0.f);
    for (size_t p = 0; p < n; ++p) {
        const double fx = x[p] * e->inv, fy = y[p] * e->inv;
        double ci = std::floor(fx), cj = std::floor(fy);
        if (!std::isfinite(ci)) ci = 0;
        if (!std::isfinite(cj)) cj = 0;
        ci = std::min(std::max(ci, 0.0), (double)(e->nx - 1));
        cj = std::min(std::max(cj, 0.0), (double)(e->ny - 1));
        float a = (float)(fx - ci), b = (float)(fy - cj);
        int ii = (int)ci, jj = (int)cj;
        if (a >= 1.f) {
            a = 0.f;
            ii = std::min(ii + 1, e->nx - 1);
        }
        if (b >= 1.f) {
            b = 0.f;
            jj = std::min(jj + 1, e->ny - 1);
        }
        if (!(a >= 0.f)) a = 0.f;
        if (!(b >= 0.f)) b = 0.f;
        cell[p] = (ii << 16) | jj;
        if (ii < 1 || ii > e->nx - 3 || jj < 1 || jj > e->ny - 3) sl[p] = slot | SLOT_OUT; // outside: parked
        fo[p] = make_float2(a, b);
        fv[p] = make_float2(vx ? (float)vx[p] : 0.f, vy ? (float)vy[p] : 0.f);
        fvz[p] = vz ? (float)vz[p] : 0.f;
        fh[p] = (float)h[p];
        fm[p] = (float)(m ? m[p] : prm.rho * h[p] * A[p]); // simulation.cpp:452
        pid[p] = (int32_t)p;
    }
    // admission by active tiles: the release's particle tiles and grid tiles
    // (its particle tiles plus right / upper / diagonal neighbours), reserved
    // tile_growth times over for the flow to spread, must fit beside the
    // larger of the live slots' reservations and the last step's counts
    int64_t res_p = 0, res_g = 0;
    if (!e->single && e->tile_growth > 0) {
        std::vector<int32_t> pt, gt;
        pt.reserve(n);
        for (size_t p = 0; p < n; ++p) {
            if (sl[p] & SLOT_OUT) continue;
            const int ii = cell[p] >> 16, jj = cell[p] & 0xffff;
            const int ib = std::min(std::max(ii - (fo[p].x < 0.5f ? 1 : 0), 0), e->nx - 3);
            const int jb = std::min(std::max(jj - (fo[p].y < 0.5f ? 1 : 0), 0), e->ny - 3);
            pt.push_back((ib >> 4) * e->nty + (jb >> 4));
        }
        std::sort(pt.begin(), pt.end());
        pt.erase(std::unique(pt.begin(), pt.end()), pt.end());
        gt.reserve(4 * pt.size());
        for (int t : pt) {
            const int tx = t / e->nty, ty = t - tx * e->nty;
            for (int dx = 0; dx < 2; ++dx)
                for (int dy = 0; dy < 2; ++dy)
                    if (tx + dx < e->ntx && ty + dy < e->nty) gt.push_back((tx + dx) * e->nty + ty + dy);
        }
        std::sort(gt.begin(), gt.end());
        gt.erase(std::unique(gt.begin(), gt.end()), gt.end());
        // a scenario never holds more than the terrain's tiles
        res_p = std::min<int64_t>((int64_t)e->tile_growth * (int64_t)pt.size(), e->ntd);
        res_g = std::min<int64_t>((int64_t)e->tile_growth * (int64_t)gt.size(), e->ntd);
        int64_t sum_p = 0, sum_g = 0;
        for (int q = 0; q < e->nslots; ++q) {
            sum_p += e->res_ptiles[q];
            sum_g += e->res_gtiles[q];
        }
        const int64_t used_p = std::max<int64_t>(sum_p, e->h_c->n_ptiles);
        const int64_t used_g = std::max<int64_t>(sum_g, e->h_c->n_gtiles);
        // touched tiles keep their max maps until release: the same reservation
        const int64_t used_h = std::max<int64_t>(sum_g, e->h_c->hm_top);
        if (used_p + res_p > e->tile_cap || used_g + res_g > e->gtile_cap || used_h + res_g > e->hm_cap) {
            if (err) *err = "tile capacity exhausted";
            return -1;
        }
    }
    const PB32& P = e->h_c->parity ? e->B : e->A;
    hipStream_t s = e->stream;
    bool ok = true;
    auto up = [&](void* dst, const void* src, size_t esz = 4) {
        ok = ok && hipMemcpyAsync((char*)dst + base * esz, src, n * esz, hipMemcpyHostToDevice, s) == hipSuccess;
    };
    up(P.cell, cell.data());
    up(P.o, fo.data(), sizeof(float2));
    up(P.v, fv.data(), sizeof(float2));
    up(P.vz, fvz.data());
    up(P.h, fh.data());
    up(P.m, fm.data());
    up(P.z, zz.data());
    up(P.pid, pid.data());
    up(P.slot, sl.data());
    // per-slot grid state starts empty
    if (e->single) // a batch slot's pool tiles were zeroed when they were returned
        ok = ok && hipMemsetAsync(e->hmax + (size_t)slot * e->nn, 0, e->nn * 4, s) == hipSuccess &&
             hipMemsetAsync(e->vmax2 + (size_t)slot * e->nn, 0, e->nn * 4, s) == hipSuccess;
    if (!ok || !hip_ok(hipStreamSynchronize(s), "admit upload", err)) {
        if (err && err->empty()) *err = "admit upload failed";
        return -1;
    }
    Slot S = prm;
    S.status = SLOT_PRIMING;
    S.time = 0.0;
    S.dt = 0.0;
    S.frame = 0;
    S.step = 0;
    S.frame_hit = 0;
    S.np = (int32_t)np;
    S.err = 0;
    e->h_slots[slot] = S;
    e->h_c->n_store = (int32_t)(base + np);
    if (e->h_c->stop == STOP_DONE) e->h_c->stop = RUNNING;
    if (!push_all(e, err)) return -1;
    const unsigned nb = (unsigned)((n + WG - 1) / WG);
    k_ez<<<nb, WG, 0, s>>>(e->d_c, e->A, e->B, (int)base, (int)(base + np), e->tz);
    k_ekeys<<<nb, WG, 0, s>>>(e->d_c, e->d_slots, e->A, e->B, (int)base, (int)(base + np), e->key, e->kcell,
                              e->tile_count);
    if (!hip_ok(hipGetLastError(), "admit kernels", err) || !hip_ok(hipStreamSynchronize(s), "admit", err)) return -1;
    e->admit_base[slot] = base;
    e->res_ptiles[slot] = res_p;
    e->res_gtiles[slot] = res_g;
    return slot;
}

bool engine_step(Engine32* e, int max_steps, std::string* err) {
    if (engine_stopped(e, err)) return false;
    hipStream_t s = e->stream;
    size_particle_grids(e);
    k_eclear<<<1, WG, 0, s>>>(e->d_c, e->d_slots);
    if (e->use_graph && !e->timing && !e->gexec && !build_graph(e, err)) return false;
    const int chunk_max = 16;
    int done = 0;
    while (done < max_steps) {
        const int chunk = std::min(chunk_max, max_steps - done);
        for (int k = 0; k < chunk; ++k) {
            if (e->use_graph && !e->timing) {
                if (!hip_ok(hipGraphLaunch(e->gexec, s), "graph launch", err)) return false;
            } else {
                enqueue_estep(e, done + k);
            }
        }
        enqueue_eend(e, done + chunk - 1);
        if (!hip_ok(hipGetLastError(), "step launch", err)) return false;
        if (!engine_pull(e, err)) return false;
        if (e->timing) collect(e, e->h_c->steps_in_call);
        done += chunk;
        if (e->h_c->stop != RUNNING) break;
    }
    for (int q = 0; q < e->nslots; ++q) e->admit_base[q] = -1; // storage was re-sorted
    if (e->h_c->stop == STOP_ERROR) {
        if (err)
            *err = e->h_c->err == 7   ? "active-tile capacity exceeded"
                   : e->h_c->err == 8 ? "max-map tile capacity exceeded"
                                      : "device error";
        return false;
    }
    return true;
}

bool engine_release(Engine32* e, int slot, std::string* err) {
    if (!engine_pull(e, err)) return false;
    if (engine_stopped(e, err)) return false;
    if (slot < 0 || slot >= e->nslots) {
        if (err) *err = "bad slot";
        return false;
    }
    e->h_slots[slot].status = SLOT_FREE;
    e->h_slots[slot].np = 0;
    e->res_ptiles[slot] = 0;
    e->res_gtiles[slot] = 0;
    if (!push_all(e, err)) return false;
    if (!e->single) {
        k_hm_release<<<(unsigned)std::min(e->ntd, 4096), WG, 0, e->stream>>>(e->d_c, slot, e->hm_map, e->hm_stack,
                                                                            e->hmax, e->vmax2);
        if (!hip_ok(hipGetLastError(), "release max maps", err)) return false;
        // the kernel moved the pool's top on the device: refresh the host
        // copy before the control block is written back below
        if (!engine_pull(e, err)) return false;
    }
    // Re-key every stored particle with the new statuses (the released slot's
    // particles get no key), then run one re-sort pass that only permutes:
    // storage is compact again and no slot advances.
    hipStream_t s = e->stream;
    const int n = e->h_c->n_store;
    if (n > 0) {
        if (!hip_ok(hipMemsetAsync(e->tile_count, 0, (size_t)e->ntiles * 4, s), "release", err) ||
            !hip_ok(hipMemsetAsync(e->cursor, 0, (size_t)e->ntiles * 4, s), "release", err))
            return false;
        k_ekeys<<<(unsigned)((n + WG - 1) / WG), WG, 0, s>>>(e->d_c, e->d_slots, e->A, e->B, 0, n, e->key, e->kcell,
                                                             e->tile_count);
        size_particle_grids(e); // storage may have grown by admissions since the last step
        e->h_c->compact = 1;
        e->h_c->stop = RUNNING;
        if (!hip_ok(hipMemcpyAsync(e->d_c, e->h_c, sizeof(ECtl), hipMemcpyHostToDevice, s), "release", err))
            return false;
        const bool tm = e->timing;
        e->timing = false;
        enqueue_estep(e, 0);
        enqueue_eend(e, 0);
        e->timing = tm;
        if (!hip_ok(hipGetLastError(), "release pass", err)) return false;
    }
    return engine_pull(e, err);
}

bool engine_particles(Engine32* e, int slot, double* const dst[9], std::string* err) {
    if (!engine_pull(e, err)) return false;
    const int n = e->h_slots[slot].np;
    if (n <= 0) return true;
    std::vector<void*> tl;
    int* cnt = nullptr;
    int32_t *oc = nullptr, *op = nullptr;
    float* of = nullptr;
    if (!ealloc(tl, &cnt, 1, err) || !ealloc(tl, &oc, n, err) || !ealloc(tl, &op, n, err) ||
        !ealloc(tl, &of, (size_t)n * 8, err)) {
        efree(tl);
        return false;
    }
    hipStream_t s = e->stream;
    bool ok = hipMemsetAsync(cnt, 0, 4, s) == hipSuccess;
    const unsigned nb = (unsigned)std::max(1, (e->h_c->n_store + WG - 1) / WG);
    k_ecollect<<<nb, WG, 0, s>>>(e->d_c, e->A, e->B, slot, cnt, oc, of, op, n);
    std::vector<int32_t> c(n), p(n);
    std::vector<float> f((size_t)n * 8);
    int got = 0;
    ok = ok && hipMemcpyAsync(&got, cnt, 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(c.data(), oc, n * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(p.data(), op, n * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(f.data(), of, (size_t)n * 32, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess;
    efree(tl);
    if (!ok || got != n) {
        if (err) *err = ok ? "particle count mismatch while collecting a slot" : "collect failed";
        return false;
    }
    const double rho = e->h_slots[slot].rho;
    for (int k = 0; k < n; ++k) {
        const int q = p[k];
        if (q < 0 || q >= n) continue;
        const float* g = &f[(size_t)k * 8];
        if (dst[0]) dst[0][q] = ((double)(c[k] >> 16) + (double)g[0]) * e->dx;
        if (dst[1]) dst[1][q] = ((double)(c[k] & 0xffff) + (double)g[1]) * e->dx;
        if (dst[2]) dst[2][q] = g[2];
        if (dst[3]) dst[3][q] = g[3];
        if (dst[4]) dst[4][q] = g[4];
        if (dst[5]) dst[5][q] = g[5];
        if (dst[6]) dst[6][q] = g[6];
        if (dst[7]) dst[7][q] = (double)g[7] / ((double)g[6] * rho);
        if (dst[8]) dst[8][q] = g[7];
    }
    return true;
}

bool engine_particles_noalloc(Engine32* e, int slot, double* const dst[9], std::string* err) {
    if (!engine_pull(e, err)) return false;
    const int n = e->h_slots[slot].np;
    if (n > NOALLOC_MAX) {
        if (err)
            *err = "slot holds " + std::to_string(n) + " particles, more than the " + std::to_string(NOALLOC_MAX) +
                   " the download without allocation stages";
        return false;
    }
    if (n <= 0) return true;
    hipStream_t s = e->stream;
    int* counter = reinterpret_cast<int*>(e->scratch_ptr);
    double* out = reinterpret_cast<double*>(e->scratch_ptr + NOALLOC_HEAD);
    k_noalloc_clear<<<1, 64, 0, s>>>(counter);
    const unsigned nb = (unsigned)std::max(1, (e->h_c->n_store + WG - 1) / WG);
    k_ecollect_pid<<<nb, WG, 0, s>>>(e->d_c, e->A, e->B, slot, n, e->h_slots[slot].rho, counter, out);
    int got = -1;
    bool ok = hipGetLastError() == hipSuccess &&
              hipMemcpyAsync(&got, counter, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess;
    for (int k = 0; k < 9; ++k)
        if (dst[k])
            ok = ok && hipMemcpyAsync(dst[k], out + (size_t)k * NOALLOC_MAX, (size_t)n * sizeof(double),
                                      hipMemcpyDeviceToHost, s) == hipSuccess;
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok || got != n) {
        if (err) *err = ok ? "particle count mismatch while collecting a slot" : "collect failed";
        return false;
    }
    return true;
}

bool engine_max_maps(Engine32* e, int slot, double* h_max, double* v_max2, std::string* err) {
    std::vector<float> f(e->nn);
    double* dst[] = {h_max, v_max2};
    float* src[] = {e->hmax + (size_t)slot * e->nn, e->vmax2 + (size_t)slot * e->nn};
    std::vector<void*> tl;
    if (!e->single) { // gather the slot's pool tiles into dense terrain arrays
        float *dh = nullptr, *dv = nullptr;
        if (!ealloc(tl, &dh, e->nn, err) || !ealloc(tl, &dv, e->nn, err)) {
            efree(tl);
            return false;
        }
        k_hm_dense<<<(unsigned)((e->nn + WG - 1) / WG), WG, 0, e->stream>>>(e->d_c, slot, e->hm_map, e->hmax, e->vmax2,
                                                                         dh, dv);
        src[0] = dh;
        src[1] = dv;
    }
    bool ok = true;
    for (int k = 0; k < 2 && ok; ++k) {
        if (!dst[k]) continue;
        ok = hip_ok(hipMemcpyAsync(f.data(), src[k], e->nn * 4, hipMemcpyDeviceToHost, e->stream), "max maps", err) &&
             hip_ok(hipStreamSynchronize(e->stream), "sync", err);
        if (ok)
            for (size_t q = 0; q < e->nn; ++q) dst[k][q] = f[q];
    }
    efree(tl);
    return ok;
}

bool engine_window(Engine32* e, int slot, int* i0, int* j0, int* wnx, int* wny, double* const dst[5],
                   std::string* err) {
    if (!engine_pull(e, err)) return false;
    const Slot& S = e->h_slots[slot];
    std::vector<void*> tl;
    unsigned long long* bb = nullptr;
    if (!ealloc(tl, &bb, 4, err)) return false;
    unsigned long long init[4] = {~0ull, 0ull, ~0ull, 0ull};
    hipStream_t s = e->stream;
    const int n = e->h_c->n_store;
    bool ok = hipMemcpyAsync(bb, init, sizeof(init), hipMemcpyHostToDevice, s) == hipSuccess;
    k_ebbox<<<(unsigned)std::max(1, (n + WG - 1) / WG), WG, 0, s>>>(e->d_c, e->A, e->B, slot, n, bb);
    unsigned long long hb[4];
    ok = ok && hipMemcpyAsync(hb, bb, sizeof(hb), hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess;
    efree(tl);
    if (!ok) {
        if (err) *err = "bbox failed";
        return false;
    }
    // remesh window, src/simulation.cpp:594-606
    const double eps = 1e-8 * e->dx;
    auto cl = [](double v) { return (int)std::max(-1e9, std::min(1e9, v)); };
    int iL = std::max(0, cl(ord_value(hb[0]) * e->inv - 1.0 - eps));
    int iR = std::min(e->nx - 1, cl(ord_value(hb[1]) * e->inv + 2.0 + eps));
    int jB = std::max(0, cl(ord_value(hb[2]) * e->inv - 1.0 - eps));
    int jT = std::min(e->ny - 1, cl(ord_value(hb[3]) * e->inv + 2.0 + eps));
    if (iR < iL || jT < jB) {
        if (err) *err = "particle window outside the terrain";
        return false;
    }
    *i0 = iL;
    *j0 = jB;
    *wnx = iR - iL + 1;
    *wny = jT - jB + 1;
    bool any = false;
    for (int k = 0; k < 5; ++k) any = any || dst[k];
    if (!any) return true;
    if (S.last_out_epoch != e->h_c->epoch) {
        if (err) *err = "last step kept no grid outputs";
        return false;
    }
    const size_t wn = (size_t)(*wnx) * (*wny);
    double* o = nullptr;
    if (!ealloc(tl, &o, 5 * wn, err)) return false;
    k_ewindow<<<(unsigned)((wn + WG - 1) / WG), WG, 0, s>>>(e->d_c, slot, iL, jB, *wnx, *wny, e->gstamp, e->vf, e->om,
                                                            e->oh, e->ovz, o);
    ok = hipStreamSynchronize(s) == hipSuccess;
    for (int k = 0; k < 5 && ok; ++k)
        if (dst[k]) ok = hipMemcpy(dst[k], o + k * wn, wn * 8, hipMemcpyDeviceToHost) == hipSuccess;
    efree(tl);
    if (!ok && err) *err = "window download failed";
    return ok;
}

bool engine_diag(Engine32* e, int slot, double out[8], std::string* err) {
    if (!engine_pull(e, err)) return false;
    const Slot& S = e->h_slots[slot];
    hipStream_t s = e->stream;
    const bool gridok = e->om && S.last_out_epoch == e->h_c->epoch;
    k_ediag<<<256, WG, 0, s>>>(e->d_c, e->A, e->B, slot, e->gtiles, e->vf, gridok ? e->om : nullptr, S.rho, e->dpart);
    std::vector<double> h(6 * 256);
    if (!hip_ok(hipMemcpyAsync(h.data(), e->dpart, h.size() * 8, hipMemcpyDeviceToHost, s), "diag", err) ||
        !hip_ok(hipStreamSynchronize(s), "sync", err))
        return false;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < 256; ++b)
        for (int q = 0; q < 6; ++q) a[q] += h[6 * b + q];
    const double nan = std::nan("");
    out[0] = gridok ? a[0] : nan;
    out[1] = gridok ? a[1] : nan;
    out[2] = gridok ? a[2] : nan;
    out[3] = a[4];
    out[4] = a[5];
    out[5] = S.dt;
    out[6] = nan; // max |v_p| is reduced on the device per step; not kept
    out[7] = gridok ? a[3] : nan;
    return true;
}

bool engine_set_velocity(Engine32* e, int slot, const double* vx, const double* vy, std::string* err) {
    if (!engine_pull(e, err)) return false;
    const int64_t base = e->admit_base[slot];
    const Slot& S = e->h_slots[slot];
    if (base < 0 || S.status != SLOT_PRIMING) {
        if (err) *err = "set the velocity before the first step";
        return false;
    }
    const size_t n = (size_t)S.np;
    std::vector<float2> v(n);
    for (size_t k = 0; k < n; ++k) v[k] = make_float2((float)vx[k], (float)vy[k]);
    const PB32& P = e->h_c->parity ? e->B : e->A;
    return hip_ok(hipMemcpy(P.v + base, v.data(), n * sizeof(float2), hipMemcpyHostToDevice), "velocity", err);
}

bool engine_set_time(Engine32* e, int slot, double time, int frame, std::string* err) {
    if (!engine_pull(e, err)) return false;
    e->h_slots[slot].time = time;
    e->h_slots[slot].frame = frame;
    return push_all(e, err);
}

bool engine_keep_grid(Engine32* e, int slot, int on, std::string* err) {
    if (!engine_pull(e, err)) return false;
    e->h_slots[slot].keep_grid = on ? 1 : 0;
    return push_all(e, err);
}

void engine_timing(Engine32* e, int enable, double* ms, int64_t* counts, int n) {
    for (int k = 0; k < n && k < kNumK; ++k) {
        if (ms) ms[k] = e->ms[k];
        if (counts) counts[k] = e->cnt[k];
    }
    if (enable && !e->timing) {
        std::fill(e->ms, e->ms + kNumK, 0.0);
        std::fill(e->cnt, e->cnt + kNumK, 0);
    }
    e->timing = enable != 0;
}

int engine_use_graph(Engine32* e, int on) {
    e->use_graph = on != 0;
    if (!on) drop_graph(e);
    return 0;
}

bool engine_autotune(Engine32* e, int trial_steps, int out[3], std::string* err) {
    // Candidate launch shapes are timed on the live workload: each trial runs
    // real steps of the simulation (the numbers do not depend on the shape),
    // first the storage re-sort period at the default workgroup count, then
    // the workgroup count at the best period, then the in-LDS capacity of the
    // tile sort at both.
    trial_steps = std::min(std::max(trial_steps, 2), 256);
    static const int kPeriods[] = {1, 4, 16, 64};
    static const int kPersist[] = {1024, 2048, 4096, 8192};
    hipEvent_t a = nullptr, b = nullptr;
    if (!hip_ok(hipEventCreate(&a), "event", err) || !hip_ok(hipEventCreate(&b), "event", err)) return false;
    const bool tm = e->timing;
    e->timing = false;
    bool ok = true, finished = false;
    auto trial = [&](int k, int p, int cap, double* ms_step) {
        if (!engine_set_reorder(e, k, err)) return false;
        e->persist = p;
        e->sort_cap = cap;
        drop_graph(e);
        if (!engine_step(e, 1, err)) return false; // captures the graph, warms caches
        // at least two re-sort periods, so a trial pays for its re-sorts and
        // for the storage order drifting between them
        const int want = std::max(trial_steps, 2 * k);
        double ms_sum = 0.0;
        int steps = 0;
        for (int rep = 0; rep < 8 && steps < want && e->h_c->stop != STOP_DONE; ++rep) {
            if (!hip_ok(hipEventRecord(a, e->stream), "event", err)) return false;
            if (!engine_step(e, want - steps, err)) return false;
            if (!hip_ok(hipEventRecord(b, e->stream), "event", err) || !hip_ok(hipEventSynchronize(b), "event", err))
                return false;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            ms_sum += ms;
            steps += e->h_c->steps_in_call;
        }
        if (steps == 0) {
            finished = true;
            *ms_step = 1e30;
        } else {
            *ms_step = ms_sum / steps;
        }
        return true;
    };
    int best_k = e->h_c->reorder_every, best_p = e->persist, best_c = e->sort_cap;
    const int cap0 = e->sort_cap;
    double best = 1e30;
    for (int k : kPeriods) {
        double ms = 0;
        if (!(ok = trial(k, kPersistE, cap0, &ms)) || finished) break;
        if (ms < best) {
            best = ms;
            best_k = k;
            best_p = kPersistE;
        }
    }
    if (ok && !finished) {
        for (int p : kPersist) {
            if (p == kPersistE) continue;
            double ms = 0;
            if (!(ok = trial(best_k, p, cap0, &ms)) || finished) break;
            if (ms < best) {
                best = ms;
                best_p = p;
            }
        }
    }
    if (ok && !finished) {
        for (int cap : ES_CAPS) {
            if (cap == cap0) continue;
            double ms = 0;
            if (!(ok = trial(best_k, best_p, cap, &ms)) || finished) break;
            if (ms < best) {
                best = ms;
                best_c = cap;
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->timing = tm;
    if (!ok) return false;
    e->persist = best_p;
    e->sort_cap = best_c;
    drop_graph(e);
    out[0] = best_k;
    out[1] = best_p;
    out[2] = best_c;
    return engine_set_reorder(e, best_k, err);
}

bool engine_set_sort_capacity(Engine32* e, int cap, std::string* err) {
    for (int c : ES_CAPS)
        if (c == cap) {
            e->sort_cap = cap;
            drop_graph(e);
            return true;
        }
    if (err) *err = "sort capacity must be 1024, 2048 or 4096 particles";
    return false;
}

void engine_launch_shape(const Engine32* e, int out[3]) {
    out[0] = e->h_c->reorder_every;
    out[1] = e->persist;
    out[2] = e->sort_cap;
}

bool engine_set_reorder(Engine32* e, int k, std::string* err) {
    if (k < 1 || k > (1 << 20)) {
        if (err) *err = "re-sort period must be in [1, 2^20]";
        return false;
    }
    if (!engine_pull(e, err)) return false;
    e->h_c->reorder_every = k;
    return push_all(e, err);
}

} // namespace slab
