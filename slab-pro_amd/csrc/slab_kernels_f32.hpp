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

// Node-independent P2G terms in sorted order; fx, fy = position relative to
// the stencil base node, in cell units (in [0.5, 1.5) away from the edges).
struct Pay32 {
    float *fx, *fy, *m, *vmx, *vmy, *vzm, *hm, *mbx, *mby, *ps;
};

__global__ void k_keys32(Ctl* c, PBuf32 A, PBuf32 B, int* key, int* tslot, int* tile_count, unsigned long long* part);
__global__ void k_init_z32(Ctl* c, PBuf32 A, const float* tz);
__global__ void k_payload32(Ctl* c, PBuf32 A, PBuf32 B, const int* idx, const float* tgx, const float* tgy,
                            const float* tcth, Pay32 Y);
__global__ void k_p2g_grid32(Ctl* c, Pay32 Y, const int* gtiles, const int* pstamp, const int* cell_begin,
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
