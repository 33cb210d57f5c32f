// slab_kernels_f32.hpp -- declarations of the FP32-form kernels.
#pragma once
#include "slab_common.hpp"

namespace slab {

// FP32 particle state: floor cell (ci*ny + cj) + in-cell offset in cell units.
struct PBuf32 {
    int32_t* cell;
    float *ox, *oy, *vx, *vy, *vz, *h, *m, *z;
    int32_t* pid;
};

// P2G partial sums per particle tile: 7 fields over the tile's 16x16 nodes
// plus the 2-node halo its stencils reach (18x18 region).
constexpr int REGION = TILE + 2;
constexpr int REGION_NODES = REGION * REGION;
constexpr int P2G_FIELDS = 7; // m, P_x, P_y, P_z, H, F_x, F_y

__global__ void k_keys32(Ctl* c, PBuf32 A, PBuf32 B, int* key, int* tslot, int* tile_count, unsigned long long* part);
__global__ void k_init_z32(Ctl* c, PBuf32 A, const float* tz);
__global__ void k_p2g_cells32(Ctl* c, PBuf32 A, PBuf32 B, const int* idx, const int* ptiles, const int* cell_begin,
                              const float* tgx, const float* tgy, const float* tcth, float* tpart);
__global__ void k_grid32(Ctl* c, const int* gtiles, const int* pstamp, const int* tile_pos, const float* tpart,
                         const float* tgx, const float* tgy, const float* tcth, float* h_max, float* v_max2,
                         float4* vf, float* out_m, float* out_h, float* out_vz);
__global__ void k_g2p32(Ctl* c, PBuf32 A, PBuf32 B, const int* idx, const float4* vf, const float* tz,
                        const float* tgx, const float* tgy, int* key, int* tslot, int* tile_count,
                        unsigned long long* part);
__global__ void k_window32(const Ctl* c, int i0, int j0, int wnx, int wny, const int* gstamp, const float4* vf,
                           const float* gm, const float* gh, const float* gvz, double* o_h, double* o_vx,
                           double* o_vy, double* o_vz, double* o_m);
__global__ void k_diag_grid32(const Ctl* c, const int* gtiles, const float4* vf, const float* gm, double* part);
__global__ void k_diag_particles32(const Ctl* c, PBuf32 A, PBuf32 B, double* part);

} // namespace slab
