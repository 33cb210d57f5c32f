// slab_kernels.hpp -- declarations of the FP64-form step kernels.
#pragma once
#include "slab_common.hpp"

namespace slab {

// Structure-of-arrays particle state (replaces class Particles,
// src/data_arrays.hpp:22-47; theta is not stored: it is looked up when used).
struct PBuf64 {
    double *x, *y, *vx, *vy, *vz, *h, *A, *m, *z;
    int32_t* pid; // creation index
};

__global__ void k_ctl_clear(Ctl* c);
__global__ void k_ctl_end(Ctl* c);
__global__ void k_prime_pre(Ctl* c);
__global__ void k_ctl_prime(Ctl* c, const unsigned long long* part);
__global__ void k_part_init(unsigned long long* part);
__global__ void k_keys(Ctl* c, PBuf64 A, PBuf64 B, int* key, int* tile_count, unsigned long long* part);
__global__ void k_init_z(Ctl* c, PBuf64 A, const double* tz);
__global__ void k_tile_reduce(Ctl* c, const int* tile_count, int ntiles, int* bsum, unsigned long long* part,
                              int with_begin);
__global__ void k_tile_scan(Ctl* c, const int* bsum, int nblk, int* boff);
__global__ void k_tile_offsets(Ctl* c, const int* tile_count, int ntiles, const int* boff, const int* bsum,
                               int nblk, int merged, int* tile_begin, int* ptiles, int* gtiles, int* pstamp,
                               int* gstamp, int* tile_pos);
__global__ void k_scatter(Ctl* c, const int* key, const int* tile_begin, int* cursor, int* tmp);
constexpr int kScatterPerThread = 8; // particles per thread of k_scatter
__global__ void k_tile_sort(Ctl* c, const int32_t* pidA, const int32_t* pidB, const int* key, const int* ptiles,
                            int* tile_count, int* cursor, const int* tile_begin, const int* tmp, int* idx,
                            int* cell_begin);
__global__ void k_p2g64(Ctl* c, PBuf64 A, PBuf64 B, const int* idx, const double* tgx, const double* tgy,
                        const double* tcth, const int* ptiles, const int* cell_begin, double* tpart);
__global__ void k_grid64(Ctl* c, const int* gtiles, const int* pstamp, const int* tile_pos, const double* tpart,
                         const double* tgx, const double* tgy, const double* tcth, double* h_max, double* v_max2,
                         double4* vf, double* out_m, double* out_h, double* out_vz);
constexpr size_t kP2G64PartialDoubles = 7 * 18 * 18; // per listed particle tile
__global__ void k_g2p(Ctl* c, PBuf64 A, PBuf64 B, const int* idx, const double4* vf, const double* tz,
                      const double* tgx, const double* tgy, int* key, int* tile_count, unsigned long long* part);
__global__ void k_window(const Ctl* c, int i0, int j0, int wnx, int wny, const int* gstamp, const double4* vf,
                         const double* gm, const double* gh, const double* gvz, double* o_h, double* o_vx,
                         double* o_vy, double* o_vz, double* o_m);
__global__ void k_diag_grid(const Ctl* c, const int* gtiles, const double4* vf, const double* gm, double* part);
__global__ void k_diag_particles(const Ctl* c, PBuf64 A, PBuf64 B, double* part);
__global__ void k_diag_final(const Ctl* c, const double* pg, int ng, const double* pp, int npb,
                             const unsigned long long* part, double* out);

} // namespace slab
