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
    if (kind < 0 || kind > 5 || n < 1 || iters < 1 || !ms_out) return set_error(SLAB_EINVAL, "bad probe arguments");
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
        const int pv = kind >= 4 ? PVN : PV;
        const unsigned nb = bytes_kind ? (unsigned)((n4 + (size_t)WG * pv - 1) / ((size_t)WG * pv))
                                       : (unsigned)((n + WG - 1) / WG);
        bool ok = true;
        if (kind == 0 || kind == 4) ok = palloc(l, &a, n4) && palloc(l, &b, n4);
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
        };
        launch(); // warm-up
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
    } while (false);
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : l) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (res != SLAB_OK) return set_error(res, why);
    return SLAB_OK;
}
