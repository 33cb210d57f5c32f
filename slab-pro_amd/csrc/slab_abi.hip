// slab_abi.hip -- host side of libslabpro.so: the C ABI declared in
// include/slab_abi.h.  FP64 runs drive the FP64-form kernels on one HIP stream;
// FP32 runs and batches drive the FP32 engine (slab_engine32.hip).
#include "slab_abi.h"
#include "slab_engine32.hpp"
#include "slab_error.hpp"
#include "slab_kernels.hpp"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace slab;

namespace {

thread_local std::string g_err;

slab_status fail(slab_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(SLAB_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                          \
    do {                                   \
        slab_status s_ = (expr);           \
        if (s_ != SLAB_OK) return s_;      \
    } while (0)

// Both forms: "p2g" is the per-tile register/LDS reduction into tile partials,
// "grid" merges the tile partials and updates the grid.  The FP32 form runs
// its tile sort and P2G as one launch, timed as "p2g" ("tile_sort" has none).
const char* const kKernelNames[] = {"ctl_begin", "tile_reduce", "tile_scan", "tile_offsets", "scatter",
                                    "tile_sort", "p2g",         "grid",      "g2p",          "ctl_end"};
constexpr int kNumKernels = 10;
constexpr int kChunk = 16;     // steps enqueued between host checks
constexpr int kPersist = 2048; // workgroups of the persistent tile kernels

} // namespace

struct slab_run {
    int device = 0;
    hipStream_t stream = nullptr;
    Ctl* d_ctl = nullptr;
    Ctl* h_ctl = nullptr; // pinned mirror
    int nx = 0, ny = 0, ntx = 0, nty = 0, ntiles = 0, nblk = 0;
    // up to this many scan blocks, k_tile_offsets forms the block offsets itself
    int scan_merge_max = 2048;
    size_t nn = 0;
    double dx = 0, inv_dx = 0;
    bool f32 = false;
    // terrain + grid, FP64 form
    double *tz = nullptr, *tgx = nullptr, *tgy = nullptr, *tcth = nullptr, *hmax = nullptr, *vmax2 = nullptr;
    double4* vf = nullptr;
    double *gm = nullptr, *gh = nullptr, *gvz = nullptr;
    // FP32 form: the engine (one slot), created at upload; terrain kept until then
    Engine32* eng = nullptr;
    std::vector<double> z32;
    Slot slot_prm{};
    // tiles
    int *tile_count = nullptr, *tile_begin = nullptr, *pstamp = nullptr, *gstamp = nullptr;
    int* cursor = nullptr; // per-tile next free slot of the scatter
    int *ptiles = nullptr, *gtiles = nullptr, *bsum = nullptr, *boff = nullptr, *cell_begin = nullptr;
    int* tile_pos = nullptr; // tile -> position in the particle-tile list
    double* tpart = nullptr; // P2G partial region per listed particle tile
    // particles
    int64_t np = 0;
    PBuf64 A{}, B{};
    int *key = nullptr, *idx = nullptr, *tmp = nullptr;
    unsigned long long* part = nullptr; // per-block |v|^2 max + bounding box partials
    double *diag_part = nullptr, *diag_out = nullptr;
    bool primed = false;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_free;
    struct Pending {
        int kid;
        int step;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double ms[kNumKernels] = {};
    int64_t cnt[kNumKernels] = {};
    std::vector<void*> allocs;
    std::vector<void*> part_allocs;
};

namespace {

template <typename T>
slab_status dalloc(T** p, size_t n, std::vector<void*>& list) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(SLAB_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                                     hipGetErrorString(e));
    }
    list.push_back(q);
    *p = (T*)q;
    return SLAB_OK;
}

slab_status pull_ctl(slab_run* r) {
    HIPCHK(hipMemcpyAsync(r->h_ctl, r->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return SLAB_OK;
}

slab_status push_ctl(slab_run* r) {
    HIPCHK(hipMemcpyAsync(r->d_ctl, r->h_ctl, sizeof(Ctl), hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return SLAB_OK;
}

hipEvent_t get_event(slab_run* r) {
    if (!r->ev_free.empty()) {
        hipEvent_t e = r->ev_free.back();
        r->ev_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Launch wrapper: optional event pair per kernel on the run's stream.
template <typename F>
void timed(slab_run* r, int kid, int step, F&& launch) {
    if (!r->timing) {
        launch();
        return;
    }
    hipEvent_t a = get_event(r), b = get_event(r);
    (void)hipEventRecord(a, r->stream);
    launch();
    (void)hipEventRecord(b, r->stream);
    r->pending.push_back({kid, step, a, b});
}

void collect_timing(slab_run* r, int steps_valid) {
    for (auto& p : r->pending) {
        if (p.step < steps_valid) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                r->ms[p.kid] += ms;
                r->cnt[p.kid] += 1;
            }
        }
        r->ev_free.push_back(p.a);
        r->ev_free.push_back(p.b);
    }
    r->pending.clear();
}

unsigned pblocks(const slab_run* r) { return (unsigned)((r->np + WG - 1) / WG); }
unsigned tblocks(const slab_run* r) { return (unsigned)std::min(r->ntiles, kPersist); }

void enqueue_keys(slab_run* r) {
    k_part_init<<<1, WG, 0, r->stream>>>(r->part);
    k_keys<<<pblocks(r), WG, 0, r->stream>>>(r->d_ctl, r->A, r->B, r->key, r->tile_count, r->part);
}

// Tile scan + bucketing + canonical sort + P2G terms: the particle order that
// stands in for remesh + sort_particles_by_grid_block (src/simulation.cpp:584-724).
// with_begin: the tile count reduction also starts the step (one extra
// block); the prime passes 0.
void enqueue_sort(slab_run* r, int step, int with_begin) {
    hipStream_t s = r->stream;
    timed(r, 1, step, [&] {
        k_tile_reduce<<<r->nblk + (with_begin ? 1 : 0), WG, 0, s>>>(r->d_ctl, r->tile_count, r->ntiles, r->bsum,
                                                                    r->part, with_begin);
    });
    const int merged = r->nblk <= r->scan_merge_max;
    if (!merged) timed(r, 2, step, [&] { k_tile_scan<<<1, WG, 0, s>>>(r->d_ctl, r->bsum, r->nblk, r->boff); });
    timed(r, 3, step, [&] {
        k_tile_offsets<<<r->nblk, WG, 0, s>>>(r->d_ctl, r->tile_count, r->ntiles, r->boff, r->bsum, r->nblk, merged,
                                               r->tile_begin, r->ptiles, r->gtiles, r->pstamp, r->gstamp,
                                               r->tile_pos);
    });
    timed(r, 4, step, [&] {
        const unsigned nb = (unsigned)((r->np + WG * kScatterPerThread - 1) / (WG * kScatterPerThread));
        k_scatter<<<nb, WG, 0, s>>>(r->d_ctl, r->key, r->tile_begin, r->cursor, r->tmp);
    });
    timed(r, 5, step, [&] {
        k_tile_sort<<<tblocks(r), WG, 0, s>>>(r->d_ctl, r->A.pid, r->B.pid, r->key, r->ptiles, r->tile_count,
                                              r->cursor, r->tile_begin, r->tmp, r->idx, r->cell_begin);
    });
}

// P2G: the per-tile register / LDS reduction into tile partials; grid:
// partial merge + update per node.
void enqueue_p2g(slab_run* r, int step) {
    hipStream_t s = r->stream;
    timed(r, 6, step, [&] {
        k_p2g64<<<tblocks(r), WG, 0, s>>>(r->d_ctl, r->A, r->B, r->idx, r->tgx, r->tgy, r->tcth, r->ptiles,
                                          r->cell_begin, r->tpart);
    });
    timed(r, 7, step, [&] {
        k_grid64<<<tblocks(r), WG, 0, s>>>(r->d_ctl, r->gtiles, r->pstamp, r->tile_pos, r->tpart, r->tgx, r->tgy,
                                           r->tcth, r->hmax, r->vmax2, r->vf, r->gm, r->gh, r->gvz);
    });
}

void enqueue_step(slab_run* r, int step) {
    hipStream_t s = r->stream;
    enqueue_sort(r, step, 1);
    enqueue_p2g(r, step);
    timed(r, 8, step, [&] {
        k_g2p<<<pblocks(r), WG, 0, s>>>(r->d_ctl, r->A, r->B, r->idx, r->vf, r->tz, r->tgx, r->tgy, r->key,
                                        r->tile_count, r->part);
    });
    timed(r, 9, step, [&] { k_ctl_end<<<1, 1, 0, s>>>(r->d_ctl); });
}

// terrainGradients + terrainTheta (src/complex_terrain.cpp:84-111) on the host
// in FP64 with the host libm, plus cos(theta) as P2G and the grid update use it
// (src/simulation.cpp:746, 825).
void terrain_tables(const double* z, int nx, int ny, double inv, std::vector<double>& gx, std::vector<double>& gy,
                    std::vector<double>& cth) {
    const size_t nn = (size_t)nx * ny;
    gx.assign(nn, 0.0);
    gy.assign(nn, 0.0);
    cth.assign(nn, 1.0);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto work = [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            for (int j = 0; j < ny; ++j) {
                const size_t k = (size_t)i * ny + j;
                if (i < nx - 1 && j < ny - 1) {
                    const double z00 = z[k], z01 = z[k + 1], z10 = z[k + ny], z11 = z[k + ny + 1];
                    gx[k] = (z10 + z11 - z00 - z01) / 2. * inv;
                    gy[k] = (z01 + z11 - z00 - z10) / 2. * inv;
                }
                const double th = std::atan(std::sqrt(gx[k] * gx[k] + gy[k] * gy[k]));
                cth[k] = std::cos(th);
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

void free_list(std::vector<void*>& l) {
    for (void* p : l) (void)hipFree(p);
    l.clear();
}

// Current-buffer particle state in creation order (FP64 form).
slab_status gather_particles(slab_run* r, double* const dst[9]) {
    TRY(pull_ctl(r));
    const size_t n = (size_t)r->np;
    const PBuf64& P = r->h_ctl->parity ? r->B : r->A;
    std::vector<int32_t> pid(n);
    HIPCHK(hipMemcpy(pid.data(), P.pid, n * 4, hipMemcpyDeviceToHost));
    std::vector<double> tmp(n);
    double* srcs[] = {P.x, P.y, P.vx, P.vy, P.vz, P.z, P.h, P.A, P.m};
    for (int f = 0; f < 9; ++f) {
        if (!dst[f]) continue;
        HIPCHK(hipMemcpy(tmp.data(), srcs[f], n * 8, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) dst[f][pid[k]] = tmp[k];
    }
    return SLAB_OK;
}

Slot slot_params(const slab_params* p, const slab_friction* f) {
    Slot S{};
    S.gravity = p->gravity;
    S.rho = p->rho;
    S.C_cfl = p->C_cfl;
    S.dt_max = p->dt_max;
    S.flip_ratio = p->flip_ratio;
    S.output_dt = p->output_dt;
    S.total_time = p->total_time;
    S.fric = f->type;
    S.mu = f->mu;
    S.xi = f->xi;
    S.mu1 = f->mu1;
    S.mu2 = f->mu2;
    S.I0 = f->I0;
    S.d = f->d;
    S.phi = f->phi;
    return S;
}

// Engine failures: an overflow of the active tiles is SLAB_EDOMAIN, every
// later call on that engine SLAB_ESTATE, anything else a device error.
slab_status eng_fail(const std::string& msg, slab_status st = SLAB_EDEVICE) {
    if (msg == kStoppedByError) st = SLAB_ESTATE;
    else if (msg.find("capacity exceeded") != std::string::npos) st = SLAB_EDOMAIN;
    return fail(st, msg);
}

// Non-finite inputs are rejected before they reach the device.
bool all_finite(const double* a, int64_t n) {
    if (!a) return true;
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(a[k])) return false;
    return true;
}

slab_status check_particles(int64_t np, const double* x, const double* y, const double* vx, const double* vy,
                            const double* vz, const double* h, const double* A, const double* m) {
    if (!all_finite(x, np) || !all_finite(y, np) || !all_finite(vx, np) || !all_finite(vy, np) ||
        !all_finite(vz, np) || !all_finite(h, np) || !all_finite(A, np) || !all_finite(m, np))
        return fail(SLAB_EINVAL, "non-finite particle input");
    return SLAB_OK;
}

slab_status check_params(const slab_params* p, const slab_friction* f) {
    const double v[] = {p->gravity, p->rho, p->C_cfl, p->dt_max, p->flip_ratio, p->output_dt, p->total_time,
                        f->mu, f->xi, f->mu1, f->mu2, f->I0, f->d, f->phi};
    if (!all_finite(v, 14)) return fail(SLAB_EINVAL, "non-finite parameter");
    return SLAB_OK;
}

} // namespace

slab_status slab::set_error(slab_status s, const std::string& msg) { return fail(s, msg); }

extern "C" {

const char* slab_last_error(void) { return g_err.c_str(); }

int32_t slab_kernel_count(void) { return kNumKernels; }

const char* slab_kernel_name(int32_t k) { return (k >= 0 && k < kNumKernels) ? kKernelNames[k] : ""; }

slab_status slab_device_info(int32_t out[3]) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return fail(SLAB_EDEVICE, "no HIP device visible");
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    out[0] = dev;
    out[1] = p.multiProcessorCount;
    out[2] = std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
    return SLAB_OK;
}

static slab_status check_inputs(const slab_terrain* terrain, const slab_params* params, const slab_friction* friction) {
    if (!terrain || !params || !friction) return fail(SLAB_EINVAL, "null argument");
    if (terrain->nx < 4 || terrain->ny < 4 || !terrain->z)
        return fail(SLAB_EINVAL, "terrain must have at least 4x4 nodes and z");
    if (!(terrain->dx > 0.0) || !std::isfinite(terrain->dx)) return fail(SLAB_EINVAL, "dx must be positive");
    if ((double)terrain->nx * (double)terrain->ny > 1.0e9) return fail(SLAB_EDOMAIN, "terrain larger than 1e9 nodes");
    if (params->precision != SLAB_FP64 && params->precision != SLAB_FP32)
        return fail(SLAB_EINVAL, "precision must be SLAB_FP64 or SLAB_FP32");
    if (friction->type < 0 || friction->type > 2) return fail(SLAB_EINVAL, "unknown friction type");
    if (!(params->output_dt > 0.0) || !(params->dt_max > 0.0))
        return fail(SLAB_EINVAL, "output_dt and dt_max must be positive");
    return check_params(params, friction);
}

slab_status slab_run_create(const slab_terrain* terrain, const slab_params* params, const slab_friction* friction,
                            slab_run** out) {
    if (!out) return fail(SLAB_EINVAL, "null argument");
    *out = nullptr;
    TRY(check_inputs(terrain, params, friction));
    const int nx = terrain->nx, ny = terrain->ny;
    int32_t info[3];
    TRY(slab_device_info(info));
    if (!info[2]) return fail(SLAB_EDEVICE, "device is not gfx950 (MI355X)");

    slab_run* r = new slab_run();
    r->device = info[0];
    r->f32 = params->precision == SLAB_FP32;
    r->nx = nx;
    r->ny = ny;
    r->nn = (size_t)nx * ny;
    r->dx = terrain->dx;
    r->inv_dx = 1. / terrain->dx;
    if (r->f32) { // the FP32 engine is sized at upload, when the particle count is known
        if (nx > 65535 || ny > 65535) {
            delete r;
            return fail(SLAB_EDOMAIN, "the FP32 form supports up to 65535 nodes per side");
        }
        r->z32.assign(terrain->z, terrain->z + r->nn);
        r->slot_prm = slot_params(params, friction);
        *out = r;
        return SLAB_OK;
    }
    r->ntx = (nx + TILE - 1) / TILE;
    r->nty = (ny + TILE - 1) / TILE;
    r->ntiles = r->ntx * r->nty;
    r->nblk = (r->ntiles + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (const char* v = std::getenv("SLAB_SCAN_MERGE_MAX_BLOCKS")) r->scan_merge_max = std::atoi(v);
    auto cleanup = [&](slab_status s) {
        slab_run_destroy(r);
        return s;
    };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(SLAB_EDEVICE, "hipStreamCreate failed"));
    if (hipHostMalloc((void**)&r->h_ctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(SLAB_ENOMEM, "hipHostMalloc failed"));
    std::memset(r->h_ctl, 0, sizeof(Ctl));
    slab_status st = SLAB_OK;
    auto A = [&](auto** p, size_t n) {
        if (st == SLAB_OK) st = dalloc(p, n, r->allocs);
    };
    A(&r->d_ctl, 1);
    A(&r->tz, r->nn);
    A(&r->tgx, r->nn);
    A(&r->tgy, r->nn);
    A(&r->tcth, r->nn);
    A(&r->hmax, r->nn);
    A(&r->vmax2, r->nn);
    A(&r->vf, r->nn);
    A(&r->gm, r->nn);
    A(&r->gh, r->nn);
    A(&r->gvz, r->nn);
    A(&r->tile_count, r->ntiles);
    A(&r->cursor, r->ntiles);
    A(&r->tile_begin, r->ntiles);
    A(&r->pstamp, r->ntiles);
    A(&r->gstamp, r->ntiles);
    A(&r->ptiles, r->ntiles);
    A(&r->gtiles, r->ntiles);
    A(&r->bsum, 3 * (size_t)r->nblk);
    A(&r->boff, 3 * (size_t)r->nblk);
    A(&r->cell_begin, (size_t)r->ntiles * CELL_STRIDE);
    A(&r->tile_pos, r->ntiles);
    A(&r->diag_part, 4 * 256 + 2 * 256);
    A(&r->diag_out, 8);
    if (st != SLAB_OK) return cleanup(st);

    std::vector<double> gx, gy, cth;
    terrain_tables(terrain->z, nx, ny, r->inv_dx, gx, gy, cth);
    hipStream_t s = r->stream;
    bool ok = true;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        ok = ok && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
    };
    auto zero = [&](void* dst, size_t bytes, int v) { ok = ok && hipMemsetAsync(dst, v, bytes, s) == hipSuccess; };
    up(r->tz, terrain->z, r->nn * 8);
    up(r->tgx, gx.data(), r->nn * 8);
    up(r->tgy, gy.data(), r->nn * 8);
    up(r->tcth, cth.data(), r->nn * 8);
    zero(r->hmax, r->nn * 8, 0);
    zero(r->vmax2, r->nn * 8, 0);
    zero(r->vf, r->nn * sizeof(double4), 0);
    zero(r->gm, r->nn * 8, 0);
    zero(r->gh, r->nn * 8, 0);
    zero(r->gvz, r->nn * 8, 0);
    zero(r->tile_count, (size_t)r->ntiles * 4, 0);
    zero(r->cursor, (size_t)r->ntiles * 4, 0);
    zero(r->pstamp, (size_t)r->ntiles * 4, 0xff);
    zero(r->gstamp, (size_t)r->ntiles * 4, 0xff);
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) return cleanup(fail(SLAB_EDEVICE, "terrain upload failed"));

    Ctl& c = *r->h_ctl;
    c.gravity = params->gravity;
    c.rho = params->rho;
    c.C_cfl = params->C_cfl;
    c.dt_max = params->dt_max;
    c.flip_ratio = params->flip_ratio;
    c.output_dt = params->output_dt;
    c.total_time = params->total_time;
    c.dx = r->dx;
    c.inv_dx = r->inv_dx;
    c.fric = friction->type;
    c.mu = friction->mu;
    c.xi = friction->xi;
    c.mu1 = friction->mu1;
    c.mu2 = friction->mu2;
    c.I0 = friction->I0;
    c.d = friction->d;
    c.phi = friction->phi;
    c.nx = nx;
    c.ny = ny;
    c.ntx = r->ntx;
    c.nty = r->nty;
    c.stop = RUNNING;
    st = push_ctl(r);
    if (st != SLAB_OK) return cleanup(st);
    *out = r;
    return SLAB_OK;
}

slab_status slab_run_upload_particles(slab_run* r, int64_t np, const double* x, const double* y, const double* vx,
                                      const double* vy, const double* vz, const double* h, const double* A,
                                      const double* m) {
    if (!r || !x || !y || !h || !A) return fail(SLAB_EINVAL, "null argument");
    if (np < 1 || np > (int64_t)1 << 30) return fail(SLAB_EINVAL, "particle count must be in [1, 2^30]");
    TRY(check_particles(np, x, y, vx, vy, vz, h, A, m));
    r->np = np;
    r->primed = false;
    if (r->f32) {
        if (r->eng) engine_destroy(r->eng);
        std::string why;
        r->eng = engine_create(r->nx, r->ny, r->dx, r->z32.data(), 1, np, true, &why);
        if (!r->eng) return fail(SLAB_ENOMEM, why);
        if (engine_admit(r->eng, r->slot_prm, np, x, y, vx, vy, vz, h, A, m, &why) != 0) return eng_fail(why);
        return SLAB_OK;
    }
    free_list(r->part_allocs);
    const size_t n = (size_t)np;
    slab_status st = SLAB_OK;
    auto P = [&](auto** q) {
        if (st == SLAB_OK) st = dalloc(q, n, r->part_allocs);
    };
    for (PBuf64* b : {&r->A, &r->B}) {
        for (double** q : {&b->x, &b->y, &b->vx, &b->vy, &b->vz, &b->h, &b->A, &b->m, &b->z}) P(q);
        P(&b->pid);
    }
    P(&r->key);
    P(&r->idx);
    P(&r->tmp);
    // P2G partials: one region per listed particle tile (at most one tile per particle)
    if (st == SLAB_OK)
        st = dalloc(&r->tpart, (size_t)std::min<int64_t>(r->ntiles, np) * kP2G64PartialDoubles, r->part_allocs);
    if (st == SLAB_OK) st = dalloc(&r->part, 5 * (size_t)PART_SHARDS, r->part_allocs);
    if (st != SLAB_OK) return st;
    TRY(pull_ctl(r));
    const double rho = r->h_ctl->rho;
    std::vector<double> mm(n), zero(n, 0.0);
    std::vector<int32_t> pid(n);
    for (size_t p = 0; p < n; ++p) {
        mm[p] = m ? m[p] : rho * h[p] * A[p]; // simulation.cpp:452
        pid[p] = (int32_t)p;
    }
    hipStream_t s = r->stream;
    const size_t nb = n * 8;
    HIPCHK(hipMemcpyAsync(r->A.x, x, nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.y, y, nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.vx, vx ? vx : zero.data(), nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.vy, vy ? vy : zero.data(), nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.vz, vz ? vz : zero.data(), nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.h, h, nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.A, A, nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.m, mm.data(), nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(r->A.pid, pid.data(), n * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    Ctl& c = *r->h_ctl;
    c.np = (int32_t)np;
    c.time = 0.0;
    c.dt = 0.0;
    c.frame = 0;
    c.step = 0;
    c.parity = 0;
    c.stop = RUNNING;
    c.err = 0;
    TRY(push_ctl(r));
    k_init_z<<<pblocks(r), WG, 0, s>>>(r->d_ctl, r->A, r->tz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return SLAB_OK;
}

slab_status slab_run_prime(slab_run* r) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (r->np < 1) return fail(SLAB_ESTATE, "no particles uploaded");
    if (r->primed) return fail(SLAB_ESTATE, "run already primed");
    if (r->f32) {
        std::string why;
        if (!engine_step(r->eng, 1, &why)) return eng_fail(why);
        if (engine_slot(r->eng, 0).status != SLOT_RUNNING) return eng_fail("priming step did not run");
        r->primed = true;
        return SLAB_OK;
    }
    hipStream_t s = r->stream;
    HIPCHK(hipMemsetAsync(r->tile_count, 0, (size_t)r->ntiles * 4, s));
    HIPCHK(hipMemsetAsync(r->cursor, 0, (size_t)r->ntiles * 4, s));
    k_prime_pre<<<1, 1, 0, s>>>(r->d_ctl);
    enqueue_keys(r);
    k_ctl_prime<<<1, WG, 0, s>>>(r->d_ctl, r->part);
    bool tm = r->timing;
    r->timing = false;
    enqueue_sort(r, 0, 0);
    enqueue_p2g(r, 0);
    r->timing = tm;
    // keys again for step 1 (the tile sort consumed the counts)
    enqueue_keys(r);
    HIPCHK(hipGetLastError());
    TRY(pull_ctl(r));
    r->primed = true;
    return SLAB_OK;
}

slab_status slab_run_step(slab_run* r, int32_t max_steps, slab_step_info* info) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (!r->primed) return fail(SLAB_ESTATE, "call slab_run_prime first");
    if (max_steps < 0) return fail(SLAB_EINVAL, "max_steps < 0");
    if (r->f32) {
        std::string why;
        if (!engine_step(r->eng, max_steps, &why)) return eng_fail(why);
        const Slot& S = engine_slot(r->eng, 0);
        if (info) {
            info->time = S.time;
            info->dt = S.dt;
            info->frame = S.frame;
            info->step = S.step;
            info->frame_hit = S.frame_hit;
            info->finished = !(S.time < S.total_time);
            info->steps_done = S.steps_in_call;
            info->reserved = 0;
        }
        if (S.status == SLOT_FAILED) return fail(SLAB_ENAN, "non-finite particle velocity on the device");
        return SLAB_OK;
    }
    hipStream_t s = r->stream;
    k_ctl_clear<<<1, 1, 0, s>>>(r->d_ctl);
    int done = 0;
    while (done < max_steps) {
        const int chunk = std::min(kChunk, max_steps - done);
        // every step of earlier chunks ran (a stop ends the loop), so done + k
        // is the call-relative index of the step enqueued here
        for (int k = 0; k < chunk; ++k) enqueue_step(r, done + k);
        HIPCHK(hipGetLastError());
        TRY(pull_ctl(r));
        if (r->timing) collect_timing(r, r->h_ctl->steps_in_call);
        done += chunk;
        if (r->h_ctl->stop != RUNNING) break;
    }
    const Ctl& c = *r->h_ctl;
    if (info) {
        info->time = c.time;
        info->dt = c.dt;
        info->frame = c.frame;
        info->step = c.step;
        info->frame_hit = c.frame_hit;
        info->finished = !(c.time < c.total_time);
        info->steps_done = c.steps_in_call;
        info->reserved = 0;
    }
    if (c.stop == STOP_ERROR) return fail(SLAB_ENAN, "non-finite particle velocity on the device");
    return SLAB_OK;
}

slab_status slab_run_download_particles(slab_run* r, double* x, double* y, double* vx, double* vy, double* vz,
                                        double* z, double* h, double* A, double* m) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (r->np < 1) return fail(SLAB_ESTATE, "no particles");
    double* const dst[9] = {x, y, vx, vy, vz, z, h, A, m};
    if (r->f32) {
        std::string why;
        return engine_particles(r->eng, 0, dst, &why) ? SLAB_OK : eng_fail(why);
    }
    return gather_particles(r, dst);
}

slab_status slab_run_keep_grid(slab_run* r, int32_t enable) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (r->f32) {
        if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
        std::string why;
        return engine_keep_grid(r->eng, 0, enable, &why) ? SLAB_OK : eng_fail(why);
    }
    TRY(pull_ctl(r));
    r->h_ctl->keep_grid = enable ? 1 : 0;
    return push_ctl(r);
}

slab_status slab_run_download_grid(slab_run* r, int32_t* i0, int32_t* j0, int32_t* wnx, int32_t* wny, double* h,
                                   double* vx, double* vy, double* vz, double* m) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (!r->primed) return fail(SLAB_ESTATE, "run not primed");
    if (r->f32) {
        std::string why;
        int a = 0, b = 0, c = 0, d = 0;
        double* const dst[5] = {h, vx, vy, vz, m};
        if (!engine_window(r->eng, 0, &a, &b, &c, &d, dst, &why))
            return fail(why.find("kept no grid") != std::string::npos ? SLAB_ESTATE : SLAB_EDEVICE, why);
        if (i0) *i0 = a;
        if (j0) *j0 = b;
        if (wnx) *wnx = c;
        if (wny) *wny = d;
        return SLAB_OK;
    }
    TRY(pull_ctl(r));
    const Ctl& c = *r->h_ctl;
    if (c.last_out_epoch != c.epoch)
        return fail(SLAB_ESTATE, "last step kept no grid outputs (not an output frame; see slab_run_keep_grid)");
    // remesh window, src/simulation.cpp:594-606 (terrain.x[0] = terrain.y[0] = 0)
    const double min_x = ord_value(c.bb_cur[0]), max_x = ord_value(c.bb_cur[1]);
    const double min_y = ord_value(c.bb_cur[2]), max_y = ord_value(c.bb_cur[3]);
    const double eps = 1e-8 * r->dx, inv = r->inv_dx;
    auto clampi = [](double v) { return (int)std::max(-1e9, std::min(1e9, v)); };
    int iL = clampi((min_x - 0.0) * inv - 1.0 - eps);
    int iR = clampi((max_x - 0.0) * inv + 2.0 + eps);
    int jB = clampi((min_y - 0.0) * inv - 1.0 - eps);
    int jT = clampi((max_y - 0.0) * inv + 2.0 + eps);
    iL = std::max(iL, 0);
    jB = std::max(jB, 0);
    iR = std::min(iR, r->nx - 1);
    jT = std::min(jT, r->ny - 1);
    if (iR < iL || jT < jB) return fail(SLAB_EDOMAIN, "particle window outside the terrain");
    const int gnx = iR - iL + 1, gny = jT - jB + 1;
    if (i0) *i0 = iL;
    if (j0) *j0 = jB;
    if (wnx) *wnx = gnx;
    if (wny) *wny = gny;
    if (!h && !vx && !vy && !vz && !m) return SLAB_OK;
    const size_t wn = (size_t)gnx * gny;
    std::vector<void*> tmpl;
    double* o = nullptr;
    slab_status st = dalloc(&o, 5 * wn, tmpl);
    if (st != SLAB_OK) return st;
    const unsigned nb = (unsigned)((wn + WG - 1) / WG);
    k_window<<<nb, WG, 0, r->stream>>>(r->d_ctl, iL, jB, gnx, gny, r->gstamp, r->vf, r->gm, r->gh, r->gvz, o, o + wn,
                                       o + 2 * wn, o + 3 * wn, o + 4 * wn);
    hipError_t e = hipStreamSynchronize(r->stream);
    double* dst[] = {h, vx, vy, vz, m};
    for (int f = 0; f < 5 && e == hipSuccess; ++f)
        if (dst[f]) e = hipMemcpy(dst[f], o + f * wn, wn * 8, hipMemcpyDeviceToHost);
    free_list(tmpl);
    if (e != hipSuccess) return fail(SLAB_EDEVICE, std::string("window download: ") + hipGetErrorString(e));
    return SLAB_OK;
}

slab_status slab_run_download_max(slab_run* r, double* h_max, double* v_max2) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (r->f32) {
        if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
        std::string why;
        return engine_max_maps(r->eng, 0, h_max, v_max2, &why) ? SLAB_OK : eng_fail(why);
    }
    HIPCHK(hipStreamSynchronize(r->stream));
    if (h_max) HIPCHK(hipMemcpy(h_max, r->hmax, r->nn * 8, hipMemcpyDeviceToHost));
    if (v_max2) HIPCHK(hipMemcpy(v_max2, r->vmax2, r->nn * 8, hipMemcpyDeviceToHost));
    return SLAB_OK;
}

slab_status slab_run_diag(slab_run* r, double out[8]) {
    if (!r || !out) return fail(SLAB_EINVAL, "null argument");
    if (!r->primed) return fail(SLAB_ESTATE, "run not primed");
    if (r->f32) {
        std::string why;
        return engine_diag(r->eng, 0, out, &why) ? SLAB_OK : eng_fail(why);
    }
    TRY(pull_ctl(r));
    hipStream_t s = r->stream;
    k_diag_grid<<<256, WG, 0, s>>>(r->d_ctl, r->gtiles, r->vf, r->gm, r->diag_part);
    k_diag_particles<<<256, WG, 0, s>>>(r->d_ctl, r->A, r->B, r->diag_part + 4 * 256);
    k_diag_final<<<1, WG, 0, s>>>(r->d_ctl, r->diag_part, 256, r->diag_part + 4 * 256, 256, r->part, r->diag_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, r->diag_out, 8 * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const Ctl& c = *r->h_ctl;
    out[5] = c.dt;
    if (c.last_out_epoch != c.epoch) out[0] = out[1] = out[2] = out[7] = std::nan("");
    return SLAB_OK;
}

slab_status slab_run_timing(slab_run* r, int32_t enable, double* ms, int64_t* counts, int32_t n) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (r->f32) {
        if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
        engine_timing(r->eng, enable, ms, counts, n);
        return SLAB_OK;
    }
    if (ms || counts) {
        for (int k = 0; k < n && k < kNumKernels; ++k) {
            if (ms) ms[k] = r->ms[k];
            if (counts) counts[k] = r->cnt[k];
        }
    }
    if (enable && !r->timing) {
        std::fill(r->ms, r->ms + kNumKernels, 0.0);
        std::fill(r->cnt, r->cnt + kNumKernels, 0);
    }
    r->timing = enable != 0;
    return SLAB_OK;
}

slab_status slab_run_set_velocity(slab_run* r, const double* vx, const double* vy) {
    if (!r || !vx || !vy) return fail(SLAB_EINVAL, "null argument");
    if (r->primed) return fail(SLAB_ESTATE, "set the velocity before slab_run_prime");
    if (r->np < 1) return fail(SLAB_ESTATE, "no particles");
    if (r->f32) {
        std::string why;
        return engine_set_velocity(r->eng, 0, vx, vy, &why) ? SLAB_OK : fail(SLAB_ESTATE, why);
    }
    TRY(pull_ctl(r));
    const bool par = r->h_ctl->parity != 0;
    const size_t n = (size_t)r->np;
    std::vector<int32_t> pid(n);
    const PBuf64& P = par ? r->B : r->A;
    HIPCHK(hipMemcpy(pid.data(), P.pid, n * 4, hipMemcpyDeviceToHost));
    std::vector<double> a(n), b(n);
    for (size_t k = 0; k < n; ++k) {
        a[k] = vx[pid[k]];
        b[k] = vy[pid[k]];
    }
    HIPCHK(hipMemcpy(P.vx, a.data(), n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.vy, b.data(), n * 8, hipMemcpyHostToDevice));
    return SLAB_OK;
}

slab_status slab_run_set_time(slab_run* r, double time, int32_t frame) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (r->f32) {
        if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
        std::string why;
        return engine_set_time(r->eng, 0, time, frame, &why) ? SLAB_OK : eng_fail(why);
    }
    TRY(pull_ctl(r));
    r->h_ctl->time = time;
    r->h_ctl->frame = frame;
    return push_ctl(r);
}

void slab_run_destroy(slab_run* r) {
    if (!r) return;
    if (r->eng) engine_destroy(r->eng);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto& p : r->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : r->ev_free) (void)hipEventDestroy(e);
    free_list(r->part_allocs);
    free_list(r->allocs);
    if (r->h_ctl) (void)hipHostFree(r->h_ctl);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

// ---------------------------------------------------------------- batches

struct slab_batch {
    Engine32* eng = nullptr;
    int nslots = 0;
};

slab_status slab_batch_create(const slab_terrain* terrain, int32_t max_slots, int64_t max_particles,
                              slab_batch** out) {
    if (!terrain || !out || !terrain->z) return fail(SLAB_EINVAL, "null argument");
    *out = nullptr;
    if (max_slots < 1 || max_slots > 65536) return fail(SLAB_EINVAL, "max_slots must be in [1, 65536]");
    if (max_particles < 1 || max_particles > ((int64_t)1 << 30))
        return fail(SLAB_EINVAL, "max_particles must be in [1, 2^30]");
    if (!(terrain->dx > 0.0)) return fail(SLAB_EINVAL, "dx must be positive");
    int32_t info[3];
    TRY(slab_device_info(info));
    if (!info[2]) return fail(SLAB_EDEVICE, "device is not gfx950 (MI355X)");
    std::string why;
    Engine32* e = engine_create(terrain->nx, terrain->ny, terrain->dx, terrain->z, max_slots, max_particles, false,
                                &why);
    if (!e) return fail(SLAB_ENOMEM, why);
    slab_batch* b = new slab_batch();
    b->eng = e;
    b->nslots = max_slots;
    *out = b;
    return SLAB_OK;
}

slab_status slab_batch_admit(slab_batch* b, const slab_params* params, const slab_friction* friction, int64_t np,
                             const double* x, const double* y, const double* vx, const double* vy, const double* h,
                             const double* A, const double* m, int32_t* slot) {
    if (!b || !params || !friction || !x || !y || !h || !A || !slot) return fail(SLAB_EINVAL, "null argument");
    *slot = -1;
    if (friction->type < 0 || friction->type > 2) return fail(SLAB_EINVAL, "unknown friction type");
    if (!(params->output_dt > 0.0) || !(params->dt_max > 0.0))
        return fail(SLAB_EINVAL, "output_dt and dt_max must be positive");
    if (np < 1) return fail(SLAB_EINVAL, "np must be >= 1");
    TRY(check_params(params, friction));
    TRY(check_particles(np, x, y, vx, vy, nullptr, h, A, m));
    std::string why;
    const int s = engine_admit(b->eng, slot_params(params, friction), np, x, y, vx, vy, nullptr, h, A, m, &why);
    if (s < 0) {
        *slot = -1;
        if (why == kStoppedByError) return fail(SLAB_ESTATE, why);
        return fail((why == "no free slot" || why == "particle capacity exhausted" || why == "tile capacity exhausted")
                        ? SLAB_ENOMEM
                        : SLAB_EDEVICE,
                    why);
    }
    *slot = s;
    return SLAB_OK;
}

slab_status slab_batch_step(slab_batch* b, int32_t max_steps, int32_t* n_active) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    if (max_steps < 0) return fail(SLAB_EINVAL, "max_steps < 0");
    std::string why;
    if (!engine_step(b->eng, max_steps, &why)) {
        if (why == kStoppedByError) return fail(SLAB_ESTATE, why);
        return fail(why.find("capacity") != std::string::npos ? SLAB_EDOMAIN : SLAB_EDEVICE, why);
    }
    if (n_active) {
        int a = 0;
        for (int s = 0; s < b->nslots; ++s) {
            const int st = engine_slot(b->eng, s).status;
            a += (st == SLOT_PRIMING || st == SLOT_RUNNING);
        }
        *n_active = a;
    }
    return SLAB_OK;
}

slab_status slab_batch_slot_info(slab_batch* b, int32_t slot, slab_slot_info* out) {
    if (!b || !out) return fail(SLAB_EINVAL, "null argument");
    if (slot < 0 || slot >= b->nslots) return fail(SLAB_EINVAL, "bad slot");
    const Slot& S = engine_slot(b->eng, slot);
    out->time = S.time;
    out->dt = S.dt;
    out->status = S.status;
    out->frame = S.frame;
    out->step = S.step;
    out->np = S.np;
    out->frame_hit = S.frame_hit;
    out->err = S.err;
    return SLAB_OK;
}

slab_status slab_batch_statuses(slab_batch* b, int32_t* status, double* time) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    for (int s = 0; s < b->nslots; ++s) {
        const Slot& S = engine_slot(b->eng, s);
        if (status) status[s] = S.status;
        if (time) time[s] = S.time;
    }
    return SLAB_OK;
}

slab_status slab_batch_download_max(slab_batch* b, int32_t slot, double* h_max, double* v_max2) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    if (slot < 0 || slot >= b->nslots) return fail(SLAB_EINVAL, "bad slot");
    std::string why;
    return engine_max_maps(b->eng, slot, h_max, v_max2, &why) ? SLAB_OK : eng_fail(why);
}

slab_status slab_batch_download_particles(slab_batch* b, int32_t slot, double* x, double* y, double* vx, double* vy,
                                          double* vz, double* z, double* h, double* A, double* m) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    if (slot < 0 || slot >= b->nslots) return fail(SLAB_EINVAL, "bad slot");
    double* const dst[9] = {x, y, vx, vy, vz, z, h, A, m};
    std::string why;
    return engine_particles(b->eng, slot, dst, &why) ? SLAB_OK : eng_fail(why);
}

slab_status slab_batch_download_particles_noalloc(slab_batch* b, int32_t slot, double* x, double* y, double* vx,
                                                  double* vy, double* vz, double* z, double* h, double* A,
                                                  double* m) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    if (slot < 0 || slot >= b->nslots) return fail(SLAB_EINVAL, "bad slot");
    double* const dst[9] = {x, y, vx, vy, vz, z, h, A, m};
    std::string why;
    if (engine_particles_noalloc(b->eng, slot, dst, &why)) return SLAB_OK;
    return fail(why.find("the download without allocation") != std::string::npos ? SLAB_EDOMAIN : SLAB_EDEVICE, why);
}

slab_status slab_batch_release(slab_batch* b, int32_t slot) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    std::string why;
    if (engine_release(b->eng, slot, &why)) return SLAB_OK;
    return fail(why == kStoppedByError ? SLAB_ESTATE : SLAB_EINVAL, why);
}

slab_status slab_batch_footprint(slab_batch* b, int64_t out[4]) {
    if (!b || !out) return fail(SLAB_EINVAL, "null argument");
    std::string why;
    if (!engine_pull(b->eng, &why)) return eng_fail(why);
    int free_slots = 0;
    for (int s = 0; s < b->nslots; ++s) free_slots += engine_slot(b->eng, s).status == SLOT_FREE;
    out[0] = engine_particle_bytes();
    out[1] = engine_slot_bytes(b->eng);
    out[2] = engine_free_particles(b->eng);
    out[3] = free_slots;
    return SLAB_OK;
}

slab_status slab_batch_tiles(slab_batch* b, int64_t out[9]) {
    if (!b || !out) return fail(SLAB_EINVAL, "null argument");
    std::string why;
    if (!engine_pull(b->eng, &why)) return eng_fail(why);
    engine_tiles(b->eng, out);
    return SLAB_OK;
}

slab_status slab_batch_set_tile_growth(slab_batch* b, int32_t factor) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    std::string why;
    return engine_set_tile_growth(b->eng, factor, &why) ? SLAB_OK : fail(SLAB_EINVAL, why);
}

slab_status slab_batch_use_graph(slab_batch* b, int32_t on) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    engine_use_graph(b->eng, on);
    return SLAB_OK;
}

slab_status slab_batch_set_resort(slab_batch* b, int32_t steps) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    std::string why;
    return engine_set_reorder(b->eng, steps, &why) ? SLAB_OK : fail(SLAB_EINVAL, why);
}

slab_status slab_run_set_resort(slab_run* r, int32_t steps) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (steps < 1) return fail(SLAB_EINVAL, "re-sort period must be >= 1");
    if (!r->f32) return SLAB_OK; // the FP64 form re-sorts its storage every step
    if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
    std::string why;
    return engine_set_reorder(r->eng, steps, &why) ? SLAB_OK : fail(SLAB_EINVAL, why);
}

slab_status slab_batch_autotune(slab_batch* b, int32_t trial_steps, int32_t out[2]) {
    if (!b || !out) return fail(SLAB_EINVAL, "null argument");
    std::string why;
    int o[3] = {0, 0, 0};
    if (!engine_autotune(b->eng, trial_steps, o, &why)) return fail(SLAB_EDEVICE, why);
    out[0] = o[0];
    out[1] = o[1];
    return SLAB_OK;
}

slab_status slab_batch_set_sort_capacity(slab_batch* b, int32_t particles) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    std::string why;
    return engine_set_sort_capacity(b->eng, particles, &why) ? SLAB_OK : fail(SLAB_EINVAL, why);
}

slab_status slab_batch_launch_shape(slab_batch* b, int32_t out[3]) {
    if (!b || !out) return fail(SLAB_EINVAL, "null argument");
    int o[3] = {0, 0, 0};
    engine_launch_shape(b->eng, o);
    for (int k = 0; k < 3; ++k) out[k] = o[k];
    return SLAB_OK;
}

slab_status slab_run_set_sort_capacity(slab_run* r, int32_t particles) {
    if (!r) return fail(SLAB_EINVAL, "null run");
    if (particles != 1024 && particles != 2048 && particles != 4096)
        return fail(SLAB_EINVAL, "sort capacity must be 1024, 2048 or 4096 particles");
    if (!r->f32) return SLAB_OK; // the FP64 form has one sort
    if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
    std::string why;
    return engine_set_sort_capacity(r->eng, particles, &why) ? SLAB_OK : fail(SLAB_EINVAL, why);
}

slab_status slab_run_launch_shape(slab_run* r, int32_t out[3]) {
    if (!r || !out) return fail(SLAB_EINVAL, "null argument");
    if (!r->f32) { // the FP64 form has one launch shape
        out[0] = 1;
        out[1] = 0;
        out[2] = 0;
        return SLAB_OK;
    }
    if (!r->eng) return fail(SLAB_ESTATE, "no particles uploaded");
    int o[3] = {0, 0, 0};
    engine_launch_shape(r->eng, o);
    for (int k = 0; k < 3; ++k) out[k] = o[k];
    return SLAB_OK;
}

slab_status slab_run_autotune(slab_run* r, int32_t trial_steps, int32_t out[2]) {
    if (!r || !out) return fail(SLAB_EINVAL, "null argument");
    if (!r->primed) return fail(SLAB_ESTATE, "call slab_run_prime first");
    if (!r->f32) { // the FP64 form has one launch shape
        out[0] = 1;
        out[1] = 0;
        return SLAB_OK;
    }
    std::string why;
    int o[3] = {0, 0, 0};
    if (!engine_autotune(r->eng, trial_steps, o, &why)) return eng_fail(why);
    out[0] = o[0];
    out[1] = o[1];
    return SLAB_OK;
}

slab_status slab_batch_timing(slab_batch* b, int32_t enable, double* ms, int64_t* counts, int32_t n) {
    if (!b) return fail(SLAB_EINVAL, "null batch");
    engine_timing(b->eng, enable, ms, counts, n);
    return SLAB_OK;
}

void slab_batch_destroy(slab_batch* b) {
    if (!b) return;
    engine_destroy(b->eng);
    delete b;
}

} // extern "C"
