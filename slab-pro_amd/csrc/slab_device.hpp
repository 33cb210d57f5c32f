// slab_device.hpp -- device helpers shared by the FP64 and FP32 forms:
// wavefront (64-lane) reductions, per-block partials of max |v|^2 and the
// particle bounding box, and run-aggregated tile counting.
#pragma once
#include "slab_common.hpp"

namespace slab {

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Data-parallel-primitive (DPP) lane moves of gfx9: VALU operations on
// registers, no LDS round trip.  Every lane of the wave must be active.
//   row_shr:k (0x110 + k): lane i of a row of 16 takes lane i - k of its row
//   row_bcast:15 (0x142) / row_bcast:31 (0x143): lane 15 / 31 to the next
//     row(s) named by the row mask
//   wave_shr:1 (0x138): lane i takes lane i - 1 of the wave
// A lane whose source does not exist, or whose row the mask leaves out,
// keeps `old` (0 here).
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned dpp_u32(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}

// Maximum of an unsigned value over the wave (every lane active), returned to
// every lane: within rows of 16 by row shifts, across rows by broadcasts to
// lane 63, then read from lane 63.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    v = max(v, dpp_u32<0x111, 0xf>(v));
    v = max(v, dpp_u32<0x112, 0xf>(v));
    v = max(v, dpp_u32<0x114, 0xf>(v));
    v = max(v, dpp_u32<0x118, 0xf>(v));
    v = max(v, dpp_u32<0x142, 0xa>(v));
    v = max(v, dpp_u32<0x143, 0xc>(v));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// The value of the previous lane of the wave (lane 0 gets 0; every lane active).
__device__ __forceinline__ int lane_prev(int v) { return (int)dpp_u32<0x138, 0xf>((unsigned)v); }

// Block-wide (256 threads = 4 waves) reduction of |v|^2 max and the bounding
// box of new positions, merged by one integer min/max atomic per field into
// shard blockIdx % PART_SHARDS (order-independent); the consumer reduces the
// shards and resets them (reset_partials).  |v|^2 >= 0, so its IEEE bits order
// like integers; NaN bits sort above +inf and are caught by ctl_begin.
__device__ __forceinline__ void block_partials(double v2, unsigned long long bx0, unsigned long long bx1, unsigned long long by0,
                               unsigned long long by1, unsigned long long* __restrict__ part) {
    __shared__ unsigned long long s_b[4][5];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long vbits = wave_max_u64((unsigned long long)__double_as_longlong(v2));
    bx0 = wave_min_u64(bx0);
    bx1 = wave_max_u64(bx1);
    by0 = wave_min_u64(by0);
    by1 = wave_max_u64(by1);
    if (lane == 0) {
        s_b[wv][0] = bx0;
        s_b[wv][1] = bx1;
        s_b[wv][2] = by0;
        s_b[wv][3] = by1;
        s_b[wv][4] = vbits;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long vmax = 0, a0 = ~0ull, a1 = 0, a2 = ~0ull, a3 = 0;
        for (int k = 0; k < 4; ++k) {
            vmax = s_b[k][4] > vmax ? s_b[k][4] : vmax;
            a0 = s_b[k][0] < a0 ? s_b[k][0] : a0;
            a1 = s_b[k][1] > a1 ? s_b[k][1] : a1;
            a2 = s_b[k][2] < a2 ? s_b[k][2] : a2;
            a3 = s_b[k][3] > a3 ? s_b[k][3] : a3;
        }
        unsigned long long* q = part + 5 * (size_t)(blockIdx.x % PART_SHARDS);
        atomicMin(&q[0], a0);
        atomicMax(&q[1], a1);
        atomicMin(&q[2], a2);
        atomicMax(&q[3], a3);
        atomicMax(&q[4], vmax);
    }
}

// Sets every shard back to the identity: ~0 for the min fields (0, 2), 0 for
// the max fields (1, 3, 4).  Threads of one block.
__device__ __forceinline__ void reset_partials(unsigned long long* __restrict__ part) {
    for (int q = threadIdx.x; q < 5 * PART_SHARDS; q += blockDim.x) {
        const int f = q % 5;
        part[q] = (f == 0 || f == 2) ? ~0ull : 0ull;
    }
}

// Reduces the partial shards (all 256 threads of a 1-block launch); result in
// out[5] on thread 0.
__device__ __forceinline__ void reduce_partials(const unsigned long long* __restrict__ part, unsigned long long out[5]) {
    __shared__ unsigned long long s_r[4][5];
    unsigned long long a0 = ~0ull, a1 = 0, a2 = ~0ull, a3 = 0, v = 0;
    const int nb = PART_SHARDS;
    for (int b = threadIdx.x; b < nb; b += WG) {
        const unsigned long long* q = part + 5 * (size_t)b;
        a0 = q[0] < a0 ? q[0] : a0;
        a1 = q[1] > a1 ? q[1] : a1;
        a2 = q[2] < a2 ? q[2] : a2;
        a3 = q[3] > a3 ? q[3] : a3;
        v = q[4] > v ? q[4] : v;
    }
    a0 = wave_min_u64(a0);
    a1 = wave_max_u64(a1);
    a2 = wave_min_u64(a2);
    a3 = wave_max_u64(a3);
    v = wave_max_u64(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_r[wv][0] = a0;
        s_r[wv][1] = a1;
        s_r[wv][2] = a2;
        s_r[wv][3] = a3;
        s_r[wv][4] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = ~0ull;
        out[1] = 0;
        out[2] = ~0ull;
        out[3] = 0;
        out[4] = 0;
        for (int k = 0; k < 4; ++k) {
            out[0] = s_r[k][0] < out[0] ? s_r[k][0] : out[0];
            out[1] = s_r[k][1] > out[1] ? s_r[k][1] : out[1];
            out[2] = s_r[k][2] < out[2] ? s_r[k][2] : out[2];
            out[3] = s_r[k][3] > out[3] ? s_r[k][3] : out[3];
            out[4] = s_r[k][4] > out[4] ? s_r[k][4] : out[4];
        }
    }
}

// Counts each lane's particle into its tile.  Lanes of a wave hold
// consecutive particles, so equal tiles form runs: one atomicAdd per run (by
// its first lane) instead of one per particle, and no value returned (the slot
// inside the tile is taken later, by the scatter).  Valid lanes must be a
// prefix of the wave; every lane of the wave must call this.
__device__ __forceinline__ void wave_tile_count(int t, bool valid, int* __restrict__ tile_count) {
    const int lane = threadIdx.x & 63;
    const int tprev = __shfl_up(t, 1, 64);
    const bool head = valid && (lane == 0 || tprev != t);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long hmask = __ballot(head);
    const int nvalid = __popcll(vmask);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hgt = hmask & ~upto;
    const int end = hgt ? (__ffsll((long long)hgt) - 1) : nvalid;
    if (head) atomicAdd(&tile_count[t], end - lane);
}

// Exclusive prefix over the scan blocks before this one, and the total, of
// the three block sums, summed by this workgroup from bsum directly (used
// when there are few scan blocks: it saves the separate scan launch).
__device__ __forceinline__ void block_prefix3(const int* __restrict__ bsum, int nblk, int pre[3], int tot[3]) {
    __shared__ int s_r[4][6];
    int v[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += WG)
        for (int k = 0; k < 3; ++k) {
            const int x = bsum[3 * b + k];
            v[k] += b < (int)blockIdx.x ? x : 0;
            v[3 + k] += x;
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 6; ++k) v[k] += __shfl_xor(v[k], o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; ++k) s_r[wv][k] = v[k];
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        pre[k] = s_r[0][k] + s_r[1][k] + s_r[2][k] + s_r[3][k];
        tot[k] = s_r[0][3 + k] + s_r[1][3 + k] + s_r[2][3 + k] + s_r[3][3 + k];
    }
    __syncthreads();
}

} // namespace slab
