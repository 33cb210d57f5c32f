// slab_kernels_f32.hip -- gfx950 kernels of the FP32 form of the MPM step.
//
// Same step structure as the FP64 form (the tile scan / bucketing / canonical
// sort kernels are shared), with FP32 state:
//   - a particle position is its floor cell (int32 linear index ci*ny + cj)
//     plus an FP32 offset (ox, oy) in [0,1) in cell units, so stencil
//     arguments u = (cell - node) + offset are exact small integers plus a
//     fraction and keep ~1e-7-cell resolution on any domain size;
//   - A is not stored: h*A = m/rho (A_init is m/(rho h) by construction,
//     src/simulation.cpp:441-452, and A is reset to m/(h rho) by G2P at :938);
//   - grid, terrain tables and max maps are FP32; simulated time, dt and the
//     conservation diagnostics stay FP64 (Ctl and k_diag_*).
// Multiply-add contraction is allowed in this translation unit.
#include "slab_common.hpp"
#include "slab_device.hpp"
#include "slab_kernels_f32.hpp"

namespace slab {

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

__device__ __forceinline__ PBuf32 cur32(const Ctl* c, const PBuf32& A, const PBuf32& B) { return c->parity ? B : A; }
__device__ __forceinline__ PBuf32 nxt32(const Ctl* c, const PBuf32& A, const PBuf32& B) { return c->parity ? A : B; }

// Stencil base of an axis from (floor cell, offset): trunc(x/dx - 0.5) =
// ci - (o < 0.5), clamped so the 3x3 stencil stays on the terrain.
__device__ __forceinline__ int base32(int ci, float o, int nmax) {
    int b = ci - (o < 0.5f ? 1 : 0);
    return b < 0 ? 0 : (b > nmax ? nmax : b);
}

__device__ __forceinline__ int key32(const Ctl* c, int ci, int cj, float ox, float oy, int* tile_out) {
    const int ib = base32(ci, ox, c->nx - 3);
    const int jb = base32(cj, oy, c->ny - 3);
    const int t = (ib >> 4) * c->nty + (jb >> 4);
    *tile_out = t;
    return t * TILE_CELLS + (((ib & 15) << 4) | (jb & 15));
}

__global__ void __launch_bounds__(WG) k_keys32(Ctl* c, PBuf32 A, PBuf32 B, int* __restrict__ key,
                                               int* __restrict__ tslot, int* __restrict__ tile_count,
                                               unsigned long long* __restrict__ part) {
    const PBuf32 P = cur32(c, A, B);
    const int s = blockIdx.x * WG + threadIdx.x;
    const bool valid = s < c->np;
    double v2 = 0.0;
    unsigned long long bx0 = ~0ull, bx1 = 0, by0 = ~0ull, by1 = 0;
    int t = 0, k = 0;
    if (valid) {
        const int cell = P.cell[s];
        const int ci = cell / c->ny, cj = cell - ci * c->ny;
        const float ox = P.ox[s], oy = P.oy[s];
        k = key32(c, ci, cj, ox, oy, &t);
        const float vx = P.vx[s], vy = P.vy[s];
        v2 = (double)(vx * vx + vy * vy);
        bx0 = bx1 = ord_bits(((double)ci + (double)ox) * c->dx);
        by0 = by1 = ord_bits(((double)cj + (double)oy) * c->dx);
    }
    const int slot = wave_tile_slot(t, valid, tile_count);
    if (valid) {
        key[s] = k;
        tslot[s] = slot;
    }
    block_partials(v2, bx0, bx1, by0, by1, part);
}

// Initial particle z, src/simulation.cpp:455-465 (4x4 quadratic weights).
__global__ void k_init_z32(Ctl* c, PBuf32 A, const float* __restrict__ tz) {
    const int s = blockIdx.x * WG + threadIdx.x;
    if (s >= c->np) return;
    const int cell = A.cell[s];
    const int ci = cell / c->ny, cj = cell - ci * c->ny;
    const float ox = A.ox[s], oy = A.oy[s];
    float z = 0.0f;
    for (int i = ci - 1; i < ci + 3; ++i) {
        for (int j = cj - 1; j < cj + 3; ++j) {
            if (i < 0 || j < 0 || i >= c->nx || j >= c->ny) continue;
            const float w = Nf((float)(ci - i) + ox) * Nf((float)(cj - j) + oy);
            z += tz[(size_t)i * c->ny + j] * w;
        }
    }
    A.z[s] = z;
}

// Quadratic B-spline weights and derivatives of the 3 stencil nodes of one
// axis for a position f in [0.5, 1.5) relative to the base node (u = f - d),
// src/tools.hpp:119-140; general branches keep the edge-clamped cases exact.
__device__ __forceinline__ void weights3f(float f, float inv, float w[3], float dw[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w[d] = Nf(f - (float)d);
        dw[d] = dNf(f - (float)d) * inv;
    }
}

// P2G, src/simulation.cpp:726-798, as a per-tile segmented reduction without
// atomics.  One workgroup per particle tile (persistent over the list), one
// thread per stencil-base cell: the thread sums its cell's particles (sorted,
// so in creation order) into 9 stencil slots x 7 fields in registers; then 9
// phases add slot (d,e) of every cell into node cell+(d,e) of an 18x18-node
// LDS region -- within a phase the cell -> node map is one-to-one, so there
// are no conflicts and the order is fixed.  The region (tile + 2-node halo) is
// written to this tile's partial; k_grid32 merges the <= 4 partials of a node.
__global__ void __launch_bounds__(WG) k_p2g_cells32(Ctl* c, PBuf32 A, PBuf32 B, const int* __restrict__ idx,
                                                    const int* __restrict__ ptiles,
                                                    const int* __restrict__ cell_begin,
                                                    const float* __restrict__ tgx, const float* __restrict__ tgy,
                                                    const float* __restrict__ tcth, float* __restrict__ tpart) {
    if (c->stop) return;
    __shared__ float s_acc[P2G_FIELDS * REGION_NODES];
    const PBuf32 P = cur32(c, A, B);
    const int nt = c->n_ptiles, ny = c->ny, nty = c->nty;
    const float g = (float)c->gravity, inv = (float)c->inv_dx;
    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = ptiles[w];
        const int tx = t / nty, ty = t - tx * nty;
        const int ib = tx * TILE + lx, jb = ty * TILE + ly;
        for (int q = threadIdx.x; q < P2G_FIELDS * REGION_NODES; q += WG) s_acc[q] = 0.f;
        float acc[9][P2G_FIELDS];
#pragma unroll
        for (int k = 0; k < 9; ++k)
#pragma unroll
            for (int f = 0; f < P2G_FIELDS; ++f) acc[k][f] = 0.f;
        const int r0 = cell_begin[(size_t)t * CELL_STRIDE + threadIdx.x];
        const int r1 = cell_begin[(size_t)t * CELL_STRIDE + threadIdx.x + 1];
        for (int r = r0; r < r1; ++r) {
            const int s = idx[r];
            const int cell = P.cell[s];
            const int ci = cell / ny, cj = cell - ci * ny;
            const float fx = (float)(ci - ib) + P.ox[s], fy = (float)(cj - jb) + P.oy[s];
            const float mp = P.m[s], hp = P.h[s];
            const float cth = tcth[cell];
            const float coeff = -mp * g * cth * cth;
            const float vmx = P.vx[s] * mp, vmy = P.vy[s] * mp, vzm = P.vz[s] * mp, hm = hp * mp;
            const float mbx = coeff * tgx[cell], mby = coeff * tgy[cell];
            // h*A * sigma_00 with h*A = m / rho and sigma_00 = -rho g cos^2 h / 2
            const float ps = -0.5f * mp * g * cth * cth * hp;
            float wx[3], wy[3], dwx[3], dwy[3];
            weights3f(fx, inv, wx, dwx);
            weights3f(fy, inv, wy, dwy);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const float wgt = wx[d] * wy[e];
                    if (wgt > 3e-3f) { // simulation.cpp:780
                        float* a = acc[d * 3 + e];
                        a[0] += mp * wgt;
                        a[1] += vmx * wgt;
                        a[2] += vmy * wgt;
                        a[3] += vzm * wgt;
                        a[4] += hm * wgt;
                        a[5] += mbx * wgt - ps * (dwx[d] * wy[e]);
                        a[6] += mby * wgt - ps * (dwy[e] * wx[d]);
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int ph = 0; ph < 9; ++ph) {
            const int node = (lx + ph / 3) * REGION + (ly + ph % 3);
#pragma unroll
            for (int f = 0; f < P2G_FIELDS; ++f) s_acc[f * REGION_NODES + node] += acc[ph][f];
            __syncthreads();
        }
        float* dst = tpart + (size_t)w * (P2G_FIELDS * REGION_NODES);
        for (int q = threadIdx.x; q < P2G_FIELDS * REGION_NODES; q += WG) dst[q] = s_acc[q];
        __syncthreads();
    }
}

// computeGridVelocity (src/simulation.cpp:800-875) for one grid tile per
// workgroup, one thread per node, after merging the node's partial sums from
// the particle tiles that reach it (own, left, below, diagonal; fixed order).
__global__ void __launch_bounds__(WG) k_grid32(Ctl* c, const int* __restrict__ gtiles, const int* __restrict__ pstamp,
                                               const int* __restrict__ tile_pos, const float* __restrict__ tpart,
                                               const float* __restrict__ tgx, const float* __restrict__ tgy,
                                               const float* __restrict__ tcth, float* __restrict__ h_max,
                                               float* __restrict__ v_max2, float4* __restrict__ vf,
                                               float* __restrict__ out_m, float* __restrict__ out_h,
                                               float* __restrict__ out_vz) {
    if (c->stop) return;
    const int nt = c->n_gtiles, ep = c->epoch, nx = c->nx, ny = c->ny, nty = c->nty;
    const float g = (float)c->gravity, dt = (float)c->dt;
    const float mu = (float)c->mu, xi = (float)c->xi, mu1 = (float)c->mu1, mu2 = (float)c->mu2;
    const float I0 = (float)c->I0, dd = (float)c->d, phi = (float)c->phi;
    const int fric = c->fric, out = c->out_step;
    const int a = threadIdx.x >> 4, b = threadIdx.x & 15;
    for (int w = blockIdx.x; w < nt; w += gridDim.x) {
        const int t = gtiles[w];
        const int tx = t / nty, ty = t - tx * nty;
        const int i = tx * TILE + a, j = ty * TILE + b;
        if (i >= nx || j >= ny) continue;
        float acc[P2G_FIELDS];
#pragma unroll
        for (int f = 0; f < P2G_FIELDS; ++f) acc[f] = 0.f;
        // (source tile, region node) pairs in fixed order
        int src[4], rn[4];
        int ns = 0;
        src[ns] = t;
        rn[ns++] = a * REGION + b;
        if (tx > 0 && a < 2) {
            src[ns] = t - nty;
            rn[ns++] = (a + TILE) * REGION + b;
        }
        if (ty > 0 && b < 2) {
            src[ns] = t - 1;
            rn[ns++] = a * REGION + (b + TILE);
        }
        if (tx > 0 && ty > 0 && a < 2 && b < 2) {
            src[ns] = t - nty - 1;
            rn[ns++] = (a + TILE) * REGION + (b + TILE);
        }
        for (int k = 0; k < ns; ++k) {
            if (pstamp[src[k]] != ep) continue;
            const float* q = tpart + (size_t)tile_pos[src[k]] * (P2G_FIELDS * REGION_NODES) + rn[k];
#pragma unroll
            for (int f = 0; f < P2G_FIELDS; ++f) acc[f] += q[f * REGION_NODES];
        }
        const float am = acc[0];
        const size_t k = (size_t)i * ny + j;
        float nvx = 0.f, nvy = 0.f, flx = 0.f, fly = 0.f, hh = 0.f, vz = 0.f;
        if (am > 1e-6f) {
            const float rm = 1.0f / am;
            const float vx = acc[1] * rm, vy = acc[2] * rm;
            vz = acc[3] * rm;
            hh = acc[4] * rm;
            const float vtx = vx + dt * (acc[5] * rm), vty = vy + dt * (acc[6] * rm);
            const float vsq = vtx * vtx + vty * vty;
            if (vsq > 1e-16f) {
                const float Gx = tgx[k], Gy = tgy[k], cth = tcth[k];
                const float v_z = vtx * Gx + vty * Gy;
                const float v3sq = vsq + v_z * v_z;
                const float vn = sqrtf(v3sq);
                float af;
                if (fric == 0) {
                    af = mu * g * cth;
                } else if (fric == 1) {
                    af = mu * g * cth + g * v3sq / (xi * hh * cth);
                } else {
                    const float I = 5.f * dd * vn / (2.f * hh * cth * cth * sqrtf(g * hh * phi));
                    const float muI = mu1 + (mu2 - mu1) / (1.f + I0 / I);
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
            if (h_max[k] < hh) h_max[k] = hh;
            const float v2 = nvx * nvx + nvy * nvy;
            if (v_max2[k] < v2) v_max2[k] = v2;
        }
        vf[k] = make_float4(nvx, nvy, flx, fly);
        if (out) {
            out_m[k] = am;
            out_h[k] = hh;
            out_vz[k] = vz;
        }
    }
}

// G2P (src/simulation.cpp:877-940) with cell + offset advection; also the next
// step's keys, tile slots and |v|^2 / bounding-box partials.
__global__ void __launch_bounds__(WG) k_g2p32(Ctl* c, PBuf32 A, PBuf32 B, const int* __restrict__ idx,
                                              const float4* __restrict__ vf, const float* __restrict__ tz,
                                              const float* __restrict__ tgx, const float* __restrict__ tgy,
                                              int* __restrict__ key, int* __restrict__ tslot,
                                              int* __restrict__ tile_count, unsigned long long* __restrict__ part) {
    if (c->stop) return;
    const PBuf32 P = cur32(c, A, B);
    const PBuf32 Q = nxt32(c, A, B);
    const int r = blockIdx.x * WG + threadIdx.x;
    const bool valid = r < c->np;
    double v2 = 0.0;
    unsigned long long bx0 = ~0ull, bx1 = 0, by0 = ~0ull, by1 = 0;
    int t = 0, kk = 0;
    if (valid) {
        const int nx = c->nx, ny = c->ny;
        const float inv = (float)c->inv_dx, dt = (float)c->dt, fr = (float)c->flip_ratio;
        const int s = idx[r];
        const int cell = P.cell[s];
        const int ci = cell / ny, cj = cell - ci * ny;
        const float ox = P.ox[s], oy = P.oy[s];
        const float Gx = tgx[cell], Gy = tgy[cell];
        const int ib = base32(ci, ox, nx - 3), jb = base32(cj, oy, ny - 3);
        const float fx = (float)(ci - ib) + ox, fy = (float)(cj - jb) + oy;
        float wx[3], wy[3], dwx[3], dwy[3];
        for (int d = 0; d < 3; ++d) {
            wx[d] = Nf(fx - (float)d);
            dwx[d] = dNf(fx - (float)d) * inv;
            wy[d] = Nf(fy - (float)d);
            dwy[d] = dNf(fy - (float)d) * inv;
        }
        float picx = 0.f, picy = 0.f, flx = 0.f, fly = 0.f, zp = 0.f, tr = 0.f;
        for (int di = 0; di < 3; ++di) {
            for (int dj = 0; dj < 3; ++dj) {
                const size_t k = (size_t)(ib + di) * ny + (jb + dj);
                const float w = wx[di] * wy[dj];
                const float4 q = vf[k];
                picx += q.x * w;
                picy += q.y * w;
                flx += q.z * w;
                fly += q.w * w;
                zp += tz[k] * w;
                tr += dwx[di] * wy[dj] * q.x + dwy[dj] * wx[di] * q.y;
            }
        }
        tr *= dt;
        const float nvx = fr * (P.vx[s] + flx) + (1.f - fr) * picx;
        const float nvy = fr * (P.vy[s] + fly) + (1.f - fr) * picy;
        // advection in cell units, then renormalise the offset into [0,1)
        float nox = ox + dt * picx * inv, noy = oy + dt * picy * inv;
        float fxs = floorf(nox), fys = floorf(noy);
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
        const float hn = P.h[s] / (1.f + tr);
        Q.cell[r] = nci * ny + ncj;
        Q.ox[r] = nox;
        Q.oy[r] = noy;
        Q.vx[r] = nvx;
        Q.vy[r] = nvy;
        Q.vz[r] = nvx * Gx + nvy * Gy;
        Q.z[r] = zp;
        Q.h[r] = hn;
        Q.m[r] = P.m[s];
        Q.pid[r] = P.pid[s];
        kk = key32(c, nci, ncj, nox, noy, &t);
        v2 = (double)(nvx * nvx + nvy * nvy);
        bx0 = bx1 = ord_bits(((double)nci + (double)nox) * c->dx);
        by0 = by1 = ord_bits(((double)ncj + (double)noy) * c->dx);
    }
    const int slot = wave_tile_slot(t, valid, tile_count);
    if (valid) {
        key[r] = kk;
        tslot[r] = slot;
    }
    block_partials(v2, bx0, bx1, by0, by1, part);
}

__global__ void k_window32(const Ctl* c, int i0, int j0, int wnx, int wny, const int* __restrict__ gstamp,
                           const float4* __restrict__ vf, const float* __restrict__ gm, const float* __restrict__ gh,
                           const float* __restrict__ gvz, double* __restrict__ o_h, double* __restrict__ o_vx,
                           double* __restrict__ o_vy, double* __restrict__ o_vz, double* __restrict__ o_m) {
    const int q = blockIdx.x * WG + threadIdx.x;
    if (q >= wnx * wny) return;
    const int a = q / wny, b = q - a * wny;
    const int i = i0 + a, j = j0 + b;
    const int t = (i >> 4) * c->nty + (j >> 4);
    double h = 0, vx = 0, vy = 0, vz = 0, m = 0;
    if (gstamp[t] == c->epoch) {
        const size_t k = (size_t)i * c->ny + j;
        const float4 v = vf[k];
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

// FP64 conservation sums of the FP32 state (same partial layout as FP64).
__global__ void __launch_bounds__(WG) k_diag_grid32(const Ctl* c, const int* __restrict__ gtiles,
                                                    const float4* __restrict__ vf, const float* __restrict__ gm,
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
            const float4 v = vf[k];
            a0 += m;
            if (m > 1e-6) {
                a1 += m * (double)v.x;
                a2 += m * (double)v.y;
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

__global__ void __launch_bounds__(WG) k_diag_particles32(const Ctl* c, PBuf32 A, PBuf32 B, double* __restrict__ part) {
    __shared__ double s[2][WG];
    const PBuf32 P = cur32(c, A, B);
    double a0 = 0, a1 = 0;
    for (int r = blockIdx.x * WG + threadIdx.x; r < c->np; r += gridDim.x * WG) {
        const double m = P.m[r];
        a0 += m;
        a1 += m / c->rho; // h*A = m / rho in this form
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

} // namespace slab
