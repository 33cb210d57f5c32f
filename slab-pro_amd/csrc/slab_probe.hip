// slab_probe.hip -- streaming kernels of known size: the device's practical
// HBM ceiling for the access patterns of the step (plain read+write copy,
// read only, write only, and the particle streams of G2P: five arrays read,
// six written in place, one element per thread as G2P walks them).
// Used by bench.py --copy-probe and its rocprof counter runs; the step's
// roofline fractions are read against the peak, these give the ceiling.
#include "slab_abi.h"
#include "slab_common.hpp"
#include "slab_error.hpp"

#include <algorithm>
#include <string>
#include <vector>

namespace slab {
namespace {

constexpr int PV = 4; // float4 per thread per pass (64 B in flight per lane)

__global__ void __launch_bounds__(WG) k_probe_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    const size_t base = ((size_t)blockIdx.x * WG * PV) + threadIdx.x;
    float4 v[PV];
#pragma unroll
    for (int k = 0; k < PV; ++k) {
        const size_t i = base + (size_t)k * WG;
        v[k] = i < n4 ? a[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < PV; ++k) {
        const size_t i = base + (size_t)k * WG;
        if (i < n4) b[i] = v[k];
    }
}

__global__ void __launch_bounds__(WG) k_probe_read(const float4* __restrict__ a, float* __restrict__ out, size_t n4) {
    const size_t base = ((size_t)blockIdx.x * WG * PV) + threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < PV; ++k) {
        const size_t i = base + (size_t)k * WG;
        if (i < n4) {
            const float4 q = a[i];
            s += (q.x + q.y) + (q.z + q.w);
        }
    }
    // one store per block keeps the loads live
    __shared__ float s_w[WG / 64];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

__global__ void __launch_bounds__(WG) k_probe_write(float4* __restrict__ b, size_t n4, float val) {
    const size_t base = ((size_t)blockIdx.x * WG * PV) + threadIdx.x;
#pragma unroll
    for (int k = 0; k < PV; ++k) {
        const size_t i = base + (size_t)k * WG;
        if (i < n4) b[i] = make_float4(val, val, val, val);
    }
}

// The same copy and write with non-temporal stores (the written lines are
// not kept in the caches) and PVN float4 per thread.
constexpr int PVN = 8;
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(WG) k_probe_copy_nt(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    const f4v* __restrict__ av = reinterpret_cast<const f4v*>(a);
    f4v* __restrict__ bv = reinterpret_cast<f4v*>(b);
    const size_t base = ((size_t)blockIdx.x * WG * PVN) + threadIdx.x;
    f4v v[PVN];
#pragma unroll
    for (int k = 0; k < PVN; ++k) {
        const size_t i = base + (size_t)k * WG;
        v[k] = i < n4 ? __builtin_nontemporal_load(&av[i]) : (f4v){0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int k = 0; k < PVN; ++k) {
        const size_t i = base + (size_t)k * WG;
        if (i < n4) __builtin_nontemporal_store(v[k], &bv[i]);
    }
}

__global__ void __launch_bounds__(WG) k_probe_write_nt(float4* __restrict__ b, size_t n4, float val) {
    f4v* __restrict__ bv = reinterpret_cast<f4v*>(b);
    const size_t base = ((size_t)blockIdx.x * WG * PVN) + threadIdx.x;
    const f4v q = {val, val, val, val};
#pragma unroll
    for (int k = 0; k < PVN; ++k) {
        const size_t i = base + (size_t)k * WG;
        if (i < n4) __builtin_nontemporal_store(q, &bv[i]);
    }
}

// G2P's particle streams without its grid gathers: read cell, (ox, oy),
// (vx, vy), h, slot (28 B); write cell, (ox, oy), (vx, vy), vz, h, z in place
// (32 B).  The arithmetic is a stand-in that keeps every value live.
__global__ void __launch_bounds__(WG) k_probe_particles(int32_t* __restrict__ cell, float2* __restrict__ o,
                                                        float2* __restrict__ v, float* __restrict__ h,
                                                        const int32_t* __restrict__ slot, float* __restrict__ vz,
                                                        float* __restrict__ z, int n) {
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= n) return;
    const int c = cell[s], sl = slot[s];
    const float2 po = o[s], pv = v[s];
    const float hp = h[s];
    const float k = (float)(sl & 7) * 1e-7f;
    cell[s] = c ^ (sl & 1);
    o[s] = make_float2(po.x + k, po.y - k);
    v[s] = make_float2(pv.x * 0.999f, pv.y * 0.999f);
    vz[s] = pv.x + pv.y;
    h[s] = hp * 0.999f;
    z[s] = po.x + hp;
}

// Phase boundaries: kPhases dependent copy phases (a -> b, then b -> a, ...,
// so each phase reads what the previous one wrote, possibly on another XCD),
// either one launch per phase (the kernel boundary orders them) or all in
// one cooperative launch whose phases are separated by grid-wide barriers
// with device-scope release / acquire fences.  Both run the same grid-stride
// copy with the same number of workgroups, so the difference of their times
// is the cost of a kernel boundary against that of a grid barrier.
constexpr int kPhases = 8;

__device__ __forceinline__ void phase_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    const size_t stride = (size_t)gridDim.x * WG;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(WG) k_probe_phase(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    phase_copy(a, b, n4);
}

// Grid-wide barrier for a cooperative launch (every workgroup resident).
// Thread 0 of each workgroup releases the workgroup's writes device-wide and
// arrives on its group's counter (workgroup index % 8, one group per XCD when
// workgroups are dealt round-robin); the last of a group arrives on the top
// counter, and the last group advances `gen`.  The others wait until `gen`
// moves.  The wait is bounded by `spin_limit` polls, after which it raises
// *fail and returns false (every workgroup then leaves the kernel).
// bar: 8 group counters 64 bytes apart, then the top counter, then gen.
constexpr int kBarGroups = 8;
constexpr int kBarStride = 16; // unsigned ints between group counters
__device__ bool grid_barrier(unsigned* bar, int* fail, unsigned long long spin_limit) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* top = bar + kBarGroups * kBarStride;
        unsigned* gen = top + kBarStride;
        const unsigned grp = blockIdx.x % kBarGroups;
        const unsigned gsize = (gridDim.x - grp + kBarGroups - 1) / kBarGroups; // workgroups in this group
        const unsigned ngrp = gridDim.x < (unsigned)kBarGroups ? gridDim.x : (unsigned)kBarGroups;
        int ok = 1;
        __threadfence();
        const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool last = false;
        if (atomicAdd(bar + grp * kBarStride, 1u) == gsize - 1) {
            atomicExch(bar + grp * kBarStride, 0u);
            if (atomicAdd(top, 1u) == ngrp - 1) {
                atomicExch(top, 0u);
                last = true;
            }
        }
        if (last) {
            __threadfence();
            atomicAdd(gen, 1u);
        } else {
            unsigned long long it = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g &&
                   __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                if (++it > spin_limit) {
                    atomicExch(fail, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            ok = __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        }
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

__global__ void __launch_bounds__(WG) k_probe_phases_coop(float4* a, float4* b, size_t n4, unsigned* bar,
                                                          int* fail) {
    for (int k = 0; k < kPhases; ++k) {
        if (k > 0 && !grid_barrier(bar, fail, 1ull << 22)) return;
        if (k & 1)
            phase_copy(b, a, n4);
        else
            phase_copy(a, b, n4);
    }
}

template <typename T>
bool palloc(std::vector<void*>& l, T** p, size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
    l.push_back(q);
    *p = (T*)q;
    return true;
}

} // namespace
} // namespace slab

using namespace slab;

extern "C" slab_status slab_stream_probe(int32_t kind, int64_t n, int32_t iters, double* ms_out) {
    if (kind < 0 || kind > 9 || n < 1 || iters < 1 || !ms_out) return set_error(SLAB_EINVAL, "bad probe arguments");
    const bool bytes_kind = kind != 3;
    if (bytes_kind && (n % 16 != 0 || n > ((int64_t)1 << 34)))
        return set_error(SLAB_EINVAL, "bytes: multiple of 16, <= 16 GiB");
    if (kind == 3 && n > ((int64_t)1 << 30)) return set_error(SLAB_EINVAL, "particles: <= 2^30");
    std::vector<void*> l;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    slab_status res = SLAB_OK;
    std::string why;
    do {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
            hipEventCreate(&e1) != hipSuccess) {
            res = SLAB_EDEVICE;
            why = "stream/event creation failed";
            break;
        }
        float4 *a = nullptr, *b = nullptr;
        float* red = nullptr;
        int32_t *cell = nullptr, *slot = nullptr;
        float2 *o = nullptr, *v = nullptr;
        float *h = nullptr, *vz = nullptr, *z = nullptr;
        const size_t n4 = bytes_kind ? (size_t)n / 16 : 0;
        const int pv = kind == 4 || kind == 5 ? PVN : PV;
        unsigned nb = bytes_kind ? (unsigned)((n4 + (size_t)WG * pv - 1) / ((size_t)WG * pv))
                                 : (unsigned)((n + WG - 1) / WG);
        unsigned* bar = nullptr;
        int* fail = nullptr;
        const bool phases = kind >= 6;
        const bool coop = kind == 7 || kind == 9;
        if (phases) {
            // phase probes: four workgroups per CU (kinds 6, 7) or one (8, 9),
            // all resident at once (the cooperative launch needs them all), the
            // same for both forms
            int dev = 0, cus = 0, per = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_probe_phases_coop),
                                                             WG, 0) != hipSuccess ||
                cus < 1 || per < 1) {
                res = SLAB_EDEVICE;
                why = "occupancy query failed";
                break;
            }
            nb = (unsigned)(cus * std::min(per, kind <= 7 ? 4 : 1));
        }
        bool ok = true;
        if (kind == 0 || kind == 4 || phases) ok = palloc(l, &a, n4) && palloc(l, &b, n4);
        if (coop) ok = ok && palloc(l, &bar, (kBarGroups + 2) * kBarStride) && palloc(l, &fail, 1);
        if (kind == 1) ok = palloc(l, &a, n4) && palloc(l, &red, nb);
        if (kind == 2 || kind == 5) ok = palloc(l, &b, n4);
        if (kind == 3)
            ok = palloc(l, &cell, n) && palloc(l, &slot, n) && palloc(l, &o, n) && palloc(l, &v, n) &&
                 palloc(l, &h, n) && palloc(l, &vz, n) && palloc(l, &z, n);
        if (!ok) {
            res = SLAB_ENOMEM;
            why = "probe allocation failed";
            break;
        }
        if (a) ok = hipMemsetAsync(a, 0, n4 * 16, st) == hipSuccess;
        if (b && phases) ok = ok && hipMemsetAsync(b, 0, n4 * 16, st) == hipSuccess;
        if (bar)
            ok = ok && hipMemsetAsync(bar, 0, (kBarGroups + 2) * kBarStride * 4, st) == hipSuccess &&
                 hipMemsetAsync(fail, 0, 4, st) == hipSuccess;
        if (kind == 3) {
            const size_t n1 = (size_t)n;
            ok = hipMemsetAsync(cell, 0, n1 * 4, st) == hipSuccess && hipMemsetAsync(slot, 0, n1 * 4, st) == hipSuccess &&
                 hipMemsetAsync(o, 0, n1 * 8, st) == hipSuccess && hipMemsetAsync(v, 0, n1 * 8, st) == hipSuccess &&
                 hipMemsetAsync(h, 0, n1 * 4, st) == hipSuccess;
        }
        if (!ok) {
            res = SLAB_EDEVICE;
            why = "probe fill failed";
            break;
        }
        auto launch = [&] {
            if (kind == 0) k_probe_copy<<<nb, WG, 0, st>>>(a, b, n4);
            if (kind == 1) k_probe_read<<<nb, WG, 0, st>>>(a, red, n4);
            if (kind == 2) k_probe_write<<<nb, WG, 0, st>>>(b, n4, 1.0f);
            if (kind == 3) k_probe_particles<<<nb, WG, 0, st>>>(cell, o, v, h, slot, vz, z, (int)n);
            if (kind == 4) k_probe_copy_nt<<<nb, WG, 0, st>>>(a, b, n4);
            if (kind == 5) k_probe_write_nt<<<nb, WG, 0, st>>>(b, n4, 1.0f);
            if (phases && !coop)
                for (int k = 0; k < kPhases; ++k) {
                    if (k & 1)
                        k_probe_phase<<<nb, WG, 0, st>>>(b, a, n4);
                    else
                        k_probe_phase<<<nb, WG, 0, st>>>(a, b, n4);
                }
            if (coop) {
                void* args[] = {&a, &b, const_cast<size_t*>(&n4), &bar, &fail};
                if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_probe_phases_coop), dim3(nb), dim3(WG),
                                               args, 0, st) != hipSuccess)
                    ok = false;
            }
        };
        launch(); // warm-up
        if (!ok) {
            (void)hipGetLastError();
            res = SLAB_EDEVICE;
            why = "cooperative launch refused";
            break;
        }
        double tot = 0.0;
        for (int it = 0; it < iters; ++it) {
            (void)hipEventRecord(e0, st);
            launch();
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                res = SLAB_EDEVICE;
                why = "probe launch failed";
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tot += ms;
        }
        *ms_out = tot / iters;
        if (fail) {
            int f = 0;
            if (hipMemcpy(&f, fail, 4, hipMemcpyDeviceToHost) != hipSuccess || f != 0) {
                res = SLAB_EDEVICE;
                why = "grid barrier wait exceeded its bound";
                break;
            }
        }
    } while (false);
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : l) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (res != SLAB_OK) return set_error(res, why);
    return SLAB_OK;
}
