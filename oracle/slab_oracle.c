/*
 * slab_oracle.c -- CPU FP64 restatement of Slab-Pro's MPM step.
 * TEST INFRASTRUCTURE ONLY (the checker); see slab_oracle.h for the scope,
 * the pinning of each function and the "parity unpinned" statement.
 * Written from the equations of the reference; every block cites the
 * reference lines (src/) whose arithmetic it restates.
 */
#define _GNU_SOURCE /* sincos */
#include "slab_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Planted mistakes, for the tests that show each closed-form check fails
 * when the term it covers is wrong (tests/test_oracle_mutants.py builds one
 * library per value; the ordinary build leaves ORC_MUTANT at 0):
 *   1 strain trace without the y term        2 strain trace without dt
 *   3 h multiplied by (1 + tr)               4 FLIP/PIC weights r and 1-r swapped
 *   5 FLIP term dropped from the blend       6 pressure without the 1/2
 *   7 pressure with cos instead of cos^2     8 pressure without h
 *   9 pressure with the opposite sign */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

struct orc_sim {
    /* terrain (Grid "terrain", data_arrays.hpp:49-93) */
    int32_t nx, ny;
    double dx, inv_dx;
    double *z, *gx, *gy, *theta, *h_max, *v_max2;
    orc_params p;
    /* particles (data_arrays.hpp:22-47) + pid = creation index for downloads */
    int32_t np;
    double *x, *y, *pz, *vx, *vy, *vz, *h, *A, *m, *th;
    int32_t* pid;
    /* window grid (remesh, simulation.cpp:584-642) */
    int32_t iL, jB, gnx, gny;
    size_t gcap;
    double *gm, *gvx, *gvy, *gvz, *gh, *gFx, *gFy, *gflx, *gfly, *gz;
    /* driver state (simulation.hpp:303-310) */
    double time, dt;
    int32_t frame, step;
    int sort_enabled;
    double t_phase[6];
};

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* tools.hpp:90-92 */
static int sgn(double v) { return (0.0 < v) - (v < 0.0); }

/* tools.hpp:119-128 */
double orc_N(double u) {
    double a = fabs(u);
    if (a < 0.5) return 0.75 - a * a;
    if (a < 1.5) return 0.5 * (1.5 - a) * (1.5 - a);
    return 0.0;
}

/* tools.hpp:131-140 */
double orc_dNdu(double u) {
    double a = fabs(u);
    if (a < 0.5) return -2.0 * u;
    if (a < 1.5) return u - 1.5 * (double)sgn(u);
    return 0.0;
}

static void* xcalloc(size_t n, size_t sz) {
    void* q = calloc(n ? n : 1, sz);
    if (!q) abort();
    return q;
}

/* complex_terrain.cpp:84-100 (gradient of the cell stored at its bottom-left
 * node; last row/column stay 0) and 102-111 (theta = atan(|G|)). */
static void terrain_fields(orc_sim* s) {
    const int32_t nx = s->nx, ny = s->ny;
    for (int32_t i = 0; i < nx - 1; ++i) {
        for (int32_t j = 0; j < ny - 1; ++j) {
            const double* z = s->z;
            double dzx = (z[(size_t)(i + 1) * ny + j] + z[(size_t)(i + 1) * ny + j + 1] - z[(size_t)i * ny + j] -
                          z[(size_t)i * ny + j + 1]) / 2. * s->inv_dx;
            double dzy = (z[(size_t)i * ny + j + 1] + z[(size_t)(i + 1) * ny + j + 1] - z[(size_t)i * ny + j] -
                          z[(size_t)(i + 1) * ny + j]) / 2. * s->inv_dx;
            s->gx[(size_t)i * ny + j] = dzx;
            s->gy[(size_t)i * ny + j] = dzy;
        }
    }
    for (size_t k = 0; k < (size_t)nx * ny; ++k) {
        double gx = s->gx[k], gy = s->gy[k];
        s->theta[k] = atan(sqrt(gx * gx + gy * gy));
    }
}

orc_sim* orc_create(int32_t nx, int32_t ny, double dx, const double* z, const orc_params* prm) {
    if (nx < 4 || ny < 4 || !(dx > 0) || !z || !prm) return NULL;
    orc_sim* s = (orc_sim*)xcalloc(1, sizeof(orc_sim));
    s->nx = nx;
    s->ny = ny;
    s->dx = dx;
    s->inv_dx = 1. / dx; /* simulation.cpp:227,231 */
    size_t n = (size_t)nx * ny;
    s->z = (double*)xcalloc(n, sizeof(double));
    memcpy(s->z, z, n * sizeof(double));
    s->gx = (double*)xcalloc(n, sizeof(double));
    s->gy = (double*)xcalloc(n, sizeof(double));
    s->theta = (double*)xcalloc(n, sizeof(double));
    s->h_max = (double*)xcalloc(n, sizeof(double));
    s->v_max2 = (double*)xcalloc(n, sizeof(double));
    s->p = *prm;
    s->sort_enabled = 1;
    terrain_fields(s);
    return s;
}

static void free_grid(orc_sim* s) {
    free(s->gm); free(s->gvx); free(s->gvy); free(s->gvz); free(s->gh);
    free(s->gFx); free(s->gFy); free(s->gflx); free(s->gfly); free(s->gz);
    s->gm = s->gvx = s->gvy = s->gvz = s->gh = s->gFx = s->gFy = s->gflx = s->gfly = s->gz = NULL;
    s->gcap = 0;
}

static void free_particles(orc_sim* s) {
    free(s->x); free(s->y); free(s->pz); free(s->vx); free(s->vy); free(s->vz);
    free(s->h); free(s->A); free(s->m); free(s->th); free(s->pid);
    s->x = s->y = s->pz = s->vx = s->vy = s->vz = s->h = s->A = s->m = s->th = NULL;
    s->pid = NULL;
}

void orc_destroy(orc_sim* s) {
    if (!s) return;
    free(s->z); free(s->gx); free(s->gy); free(s->theta); free(s->h_max); free(s->v_max2);
    free_particles(s);
    free_grid(s);
    free(s);
}

/* floor-cell lookup of particleSlopeAngle / particleGradients,
 * complex_terrain.cpp:113-130 (terrain.x[0] = terrain.y[0] = 0). */
static size_t floor_cell(const orc_sim* s, double x, double y) {
    int32_t i = (int32_t)floor((x - 0.0) * s->inv_dx);
    int32_t j = (int32_t)floor((y - 0.0) * s->inv_dx);
    if (i < 0) i = 0;               /* out-of-terrain reads are undefined in the */
    if (j < 0) j = 0;               /* reference; the clamp only keeps this      */
    if (i > s->nx - 1) i = s->nx - 1; /* checker memory-safe for such inputs     */
    if (j > s->ny - 1) j = s->ny - 1;
    return (size_t)i * s->ny + j;
}

int orc_set_particles(orc_sim* s, int32_t np, const double* x, const double* y, const double* h,
                      const double* A, const double* m) {
    if (!s || np < 0) return -1;
    free_particles(s);
    s->np = np;
    s->x = (double*)xcalloc(np, 8); s->y = (double*)xcalloc(np, 8); s->pz = (double*)xcalloc(np, 8);
    s->vx = (double*)xcalloc(np, 8); s->vy = (double*)xcalloc(np, 8); s->vz = (double*)xcalloc(np, 8);
    s->h = (double*)xcalloc(np, 8); s->A = (double*)xcalloc(np, 8); s->m = (double*)xcalloc(np, 8);
    s->th = (double*)xcalloc(np, 8); s->pid = (int32_t*)xcalloc(np, 4);
    for (int32_t p = 0; p < np; ++p) {
        s->x[p] = x[p];
        s->y[p] = y[p];
        s->h[p] = h[p];
        s->A[p] = A[p];
        /* simulation.cpp:452 */
        s->m[p] = m ? m[p] : s->p.rho * s->h[p] * s->A[p];
        s->pid[p] = p;
        s->th[p] = s->theta[floor_cell(s, x[p], y[p])]; /* 453 */
        /* simulation.cpp:455-465: 4x4 quadratic interpolation of z */
        int32_t ib = (int32_t)((x[p] - 0.0) * s->inv_dx) - 1;
        int32_t jb = (int32_t)((y[p] - 0.0) * s->inv_dx) - 1;
        double zp = 0.0;
        for (int32_t i = ib; i < ib + 4; ++i) {
            for (int32_t j = jb; j < jb + 4; ++j) {
                if (i < 0 || j < 0 || i >= s->nx || j >= s->ny) continue;
                double w = orc_N((x[p] - (double)i * s->dx) * s->inv_dx) *
                           orc_N((y[p] - (double)j * s->dx) * s->inv_dx);
                zp += s->z[(size_t)i * s->ny + j] * w;
            }
        }
        s->pz[p] = zp;
    }
    s->time = 0.0;
    s->dt = 0.0; /* simulation.hpp:39: dt starts at zero */
    s->frame = 0;
    s->step = 0;
    return 0;
}

/* simulation.cpp:565-582 */
static void update_dt(orc_sim* s) {
    double best = -1.0;
    for (int32_t p = 0; p < s->np; ++p) {
        double q = s->vx[p] * s->vx[p] + s->vy[p] * s->vy[p];
        if (best < q) best = q; /* max_element keeps the first maximum */
    }
    double max_speed = s->np > 0 ? sqrt(best) : 0.0;
    if (max_speed > 0.0) {
        double c = s->p.C_cfl * s->dx / max_speed;
        s->dt = c < s->p.dt_max ? c : s->p.dt_max; /* std::min(a,b) = b<a?b:a */
    } else {
        s->dt = s->p.dt_max;
    }
    double next_frame_time = s->p.output_dt * (double)(s->frame + 1);
    if (s->time + s->dt > next_frame_time - 1e-12) s->dt = next_frame_time - s->time;
}

/* simulation.cpp:584-642 */
static void remesh(orc_sim* s) {
    double min_x = s->x[0], max_x = s->x[0], min_y = s->y[0], max_y = s->y[0];
    for (int32_t p = 1; p < s->np; ++p) {
        if (s->x[p] < min_x) min_x = s->x[p];
        if (!(s->x[p] < max_x)) max_x = s->x[p];
        if (s->y[p] < min_y) min_y = s->y[p];
        if (!(s->y[p] < max_y)) max_y = s->y[p];
    }
    const double eps = 1e-8 * s->dx;
    int32_t iL = (int32_t)((min_x - 0.0) * s->inv_dx - 1.0 - eps);
    int32_t iR = (int32_t)((max_x - 0.0) * s->inv_dx + 2.0 + eps);
    int32_t jB = (int32_t)((min_y - 0.0) * s->inv_dx - 1.0 - eps);
    int32_t jT = (int32_t)((max_y - 0.0) * s->inv_dx + 2.0 + eps);
    /* out-of-terrain windows are undefined behaviour in the reference (634-640);
     * clamp so the checker itself stays memory-safe */
    if (iL < 0) iL = 0;
    if (jB < 0) jB = 0;
    if (iR > s->nx - 1) iR = s->nx - 1;
    if (jT > s->ny - 1) jT = s->ny - 1;
    s->iL = iL;
    s->jB = jB;
    s->gnx = iR - iL + 1;
    s->gny = jT - jB + 1;
    size_t n = (size_t)s->gnx * s->gny;
    if (n > s->gcap) {
        free_grid(s);
        s->gcap = n;
        s->gm = (double*)xcalloc(n, 8); s->gvx = (double*)xcalloc(n, 8); s->gvy = (double*)xcalloc(n, 8);
        s->gvz = (double*)xcalloc(n, 8); s->gh = (double*)xcalloc(n, 8); s->gFx = (double*)xcalloc(n, 8);
        s->gFy = (double*)xcalloc(n, 8); s->gflx = (double*)xcalloc(n, 8); s->gfly = (double*)xcalloc(n, 8);
        s->gz = (double*)xcalloc(n, 8);
    }
    memset(s->gm, 0, n * 8); memset(s->gvx, 0, n * 8); memset(s->gvy, 0, n * 8);
    memset(s->gvz, 0, n * 8); memset(s->gh, 0, n * 8); memset(s->gFx, 0, n * 8);
    memset(s->gFy, 0, n * 8); memset(s->gflx, 0, n * 8); memset(s->gfly, 0, n * 8);
    for (int32_t i = 0; i < s->gnx; ++i)
        for (int32_t j = 0; j < s->gny; ++j)
            s->gz[(size_t)i * s->gny + j] = s->z[(size_t)(i + iL) * s->ny + (j + jB)];
}

static void permute_d(double* a, const int32_t* order, int32_t n, double* tmp) {
    for (int32_t k = 0; k < n; ++k) tmp[k] = a[order[k]];
    memcpy(a, tmp, (size_t)n * 8);
}

/* simulation.cpp:644-724: counting sort by 4x4 block of the window. */
static void sort_particles(orc_sim* s) {
    const int B = 4;
    const int32_t nbx = (s->gnx + B - 1) / B, nby = (s->gny + B - 1) / B;
    const int32_t nb = nbx * nby;
    int32_t* counts = (int32_t*)xcalloc(nb, 4);
    int32_t* bid = (int32_t*)xcalloc(s->np, 4);
    int32_t* order = (int32_t*)xcalloc(s->np, 4);
    const double gx0 = (double)s->iL * s->dx, gy0 = (double)s->jB * s->dx;
    for (int32_t p = 0; p < s->np; ++p) {
        int32_t ib = (int32_t)((s->x[p] - gx0) * s->inv_dx - 0.5);
        int32_t jb = (int32_t)((s->y[p] - gy0) * s->inv_dx - 0.5);
        if (ib < 0) ib = 0;
        if (jb < 0) jb = 0;
        if (ib >= s->gnx) ib = s->gnx - 1;
        if (jb >= s->gny) jb = s->gny - 1;
        int32_t b = (ib / B) * nby + jb / B;
        bid[p] = b;
        counts[b]++;
    }
    int32_t sum = 0;
    for (int32_t b = 0; b < nb; ++b) {
        int32_t c = counts[b];
        counts[b] = sum;
        sum += c;
    }
    for (int32_t p = 0; p < s->np; ++p) order[counts[bid[p]]++] = p;
    double* tmp = (double*)xcalloc(s->np, 8);
    permute_d(s->x, order, s->np, tmp); permute_d(s->y, order, s->np, tmp);
    permute_d(s->pz, order, s->np, tmp); permute_d(s->vx, order, s->np, tmp);
    permute_d(s->vy, order, s->np, tmp); permute_d(s->vz, order, s->np, tmp);
    permute_d(s->h, order, s->np, tmp); permute_d(s->A, order, s->np, tmp);
    permute_d(s->m, order, s->np, tmp); permute_d(s->th, order, s->np, tmp);
    int32_t* tp = (int32_t*)xcalloc(s->np, 4);
    for (int32_t k = 0; k < s->np; ++k) tp[k] = s->pid[order[k]];
    memcpy(s->pid, tp, (size_t)s->np * 4);
    free(tp); free(tmp); free(order); free(bid); free(counts);
}

/* Weights of the 3 stencil nodes along one axis (simulation.cpp:759-771). */
static inline void weights3(double xp, int32_t base_global, double dx, double inv, double w[3], double dw[3]) {
    for (int d = 0; d < 3; ++d) {
        double u = (xp - (double)(base_global + d) * dx) * inv;
        w[d] = orc_N(u);
        dw[d] = orc_dNdu(u) * inv;
    }
}

/* simulation.cpp:726-798 */
static void p2g(orc_sim* s) {
    const double inv = s->inv_dx, dx = s->dx, g = s->p.gravity, rho = s->p.rho;
    const double gx0 = (double)s->iL * dx, gy0 = (double)s->jB * dx;
    const int32_t gny = s->gny;
    for (int32_t p = 0; p < s->np; ++p) {
        const double xp = s->x[p], yp = s->y[p], hp = s->h[p], Ap = s->A[p], mp = s->m[p];
        const double vmx = s->vx[p] * mp, vmy = s->vy[p] * mp;
        const double vzm = s->vz[p] * mp, hm = hp * mp, hA = hp * Ap;
        int32_t ib = (int32_t)((xp - gx0) * inv - 0.5);
        int32_t jb = (int32_t)((yp - gy0) * inv - 0.5);
        size_t fc = floor_cell(s, xp, yp);
        const double Gx = s->gx[fc], Gy = s->gy[fc];
        s->th[p] = s->theta[fc];
        const double c = cos(s->th[p]);
        const double coeff = -mp * g * c * c;
        const double mbx = coeff * Gx, mby = coeff * Gy;
#if ORC_MUTANT == 6
        const double s00 = -1.0 * rho * g * c * c * hp, s01 = 0.0, s10 = 0.0, s11 = s00;
#elif ORC_MUTANT == 7
        const double s00 = -0.5 * rho * g * c * hp, s01 = 0.0, s10 = 0.0, s11 = s00;
#elif ORC_MUTANT == 8
        const double s00 = -0.5 * rho * g * c * c, s01 = 0.0, s10 = 0.0, s11 = s00;
#elif ORC_MUTANT == 9
        const double s00 = 0.5 * rho * g * c * c * hp, s01 = 0.0, s10 = 0.0, s11 = s00;
#else
        /* simulation.cpp:753 */
        const double s00 = -0.5 * rho * g * c * c * hp, s01 = 0.0, s10 = 0.0, s11 = s00;
#endif
        double wx[3], wy[3], dwx[3], dwy[3];
        weights3(xp, ib + s->iL, dx, inv, wx, dwx);
        weights3(yp, jb + s->jB, dx, inv, wy, dwy);
        for (int di = 0; di < 3; ++di) {
            int32_t i = ib + di;
            if (i < 0 || i >= s->gnx) continue; /* memory guard; UB in the reference */
            for (int dj = 0; dj < 3; ++dj) {
                int32_t j = jb + dj;
                if (j < 0 || j >= gny) continue;
                const double w = wx[di] * wy[dj];
                if (w > 3e-3) {
                    const double gwx = dwx[di] * wy[dj];
                    const double gwy = dwy[dj] * wx[di];
                    size_t k = (size_t)i * gny + j;
                    s->gm[k] += mp * w;
                    s->gvx[k] += vmx * w;
                    s->gvy[k] += vmy * w;
                    s->gvz[k] += vzm * w;
                    s->gh[k] += hm * w;
                    s->gFx[k] += mbx * w - hA * (s00 * gwx + s01 * gwy);
                    s->gFy[k] += mby * w - hA * (s11 * gwy + s10 * gwx);
                }
            }
        }
    }
}

/* simulation.cpp:800-875 */
static void grid_update(orc_sim* s) {
    const double dt = s->dt, g = s->p.gravity;
    const orc_params* P = &s->p;
    for (int32_t i = 0; i < s->gnx; ++i) {
        for (int32_t j = 0; j < s->gny; ++j) {
            size_t k = (size_t)i * s->gny + j;
            const double mi = s->gm[k];
            if (mi > 1e-6) {
                s->gvx[k] /= mi;
                s->gvy[k] /= mi;
                s->gvz[k] /= mi;
                s->gh[k] /= mi;
                const double vox = s->gvx[k], voy = s->gvy[k];
                const double ax = s->gFx[k] / mi, ay = s->gFy[k] / mi;
                const double vtx = vox + dt * ax, vty = voy + dt * ay;
                const double vsq = vtx * vtx + vty * vty;
                size_t t = (size_t)(s->iL + i) * s->ny + (s->jB + j);
                double nvx, nvy;
                if (vsq > 1e-16) {
                    const double Gx = s->gx[t], Gy = s->gy[t];
                    const double c = cos(s->theta[t]);
                    const double v_z = vtx * Gx + vty * Gy;
                    const double v3sq = vsq + v_z * v_z;
                    const double vn = sqrt(v3sq);
                    double af = 0.0;
                    const double hh = s->gh[k];
                    switch (P->fric_type) {
                        case ORC_COULOMB: af = P->mu * g * c; break;
                        case ORC_VOELLMY: af = P->mu * g * c + g * v3sq / (P->xi * hh * c); break;
                        default: {
                            const double I = 5. * P->d * vn / (2. * hh * c * c * sqrt(g * hh * P->phi));
                            const double muI = P->mu1 + (P->mu2 - P->mu1) / (1. + P->I0 / I);
                            af = muI * g * c;
                        }
                    }
                    const double dv = dt * af;
                    if (dv < vn) {
                        const double f = 1. - dv / vn;
                        nvx = vtx * f;
                        nvy = vty * f;
                    } else {
                        nvx = 0.0;
                        nvy = 0.0;
                    }
                } else {
                    nvx = 0.0;
                    nvy = 0.0;
                }
                s->gvx[k] = nvx;
                s->gvy[k] = nvy;
                s->gflx[k] = nvx - vox;
                s->gfly[k] = nvy - voy;
                if (s->h_max[t] < s->gh[k]) s->h_max[t] = s->gh[k];
                const double v2 = nvx * nvx + nvy * nvy;
                if (s->v_max2[t] < v2) s->v_max2[t] = v2;
            } else {
                s->gvx[k] = 0.0;
                s->gvy[k] = 0.0;
                s->gvz[k] = 0.0;
                s->gh[k] = 0.0;
            }
        }
    }
}

/* simulation.cpp:877-940 */
static void g2p(orc_sim* s) {
    const double inv = s->inv_dx, dx = s->dx, dt = s->dt, r = s->p.flip_ratio, rho = s->p.rho;
    const double gx0 = (double)s->iL * dx, gy0 = (double)s->jB * dx;
    const int32_t gny = s->gny;
    for (int32_t p = 0; p < s->np; ++p) {
        const double xp = s->x[p], yp = s->y[p];
        size_t fc = floor_cell(s, xp, yp);
        const double Gx = s->gx[fc], Gy = s->gy[fc];
        int32_t ib = (int32_t)((xp - gx0) * inv - 0.5);
        int32_t jb = (int32_t)((yp - gy0) * inv - 0.5);
        double picx = 0, picy = 0, flx = 0, fly = 0, zp = 0, tr = 0;
        double wx[3], wy[3], dwx[3], dwy[3];
        weights3(xp, ib + s->iL, dx, inv, wx, dwx);
        weights3(yp, jb + s->jB, dx, inv, wy, dwy);
        for (int di = 0; di < 3; ++di) {
            int32_t i = ib + di;
            if (i < 0 || i >= s->gnx) continue;
            for (int dj = 0; dj < 3; ++dj) {
                int32_t j = jb + dj;
                if (j < 0 || j >= gny) continue;
                const double w = wx[di] * wy[dj];
                size_t k = (size_t)i * gny + j;
                const double gwx = dwx[di] * wy[dj];
                const double gwy = dwy[dj] * wx[di];
                picx += s->gvx[k] * w;
                picy += s->gvy[k] * w;
                flx += s->gflx[k] * w;
                fly += s->gfly[k] * w;
                zp += s->gz[k] * w;
#if ORC_MUTANT == 1
                tr += gwx * s->gvx[k];
#else
                tr += gwx * s->gvx[k] + gwy * s->gvy[k]; /* simulation.cpp:924-928 */
#endif
            }
        }
#if ORC_MUTANT != 2
        tr *= dt;
#endif
        s->pz[p] = zp;
#if ORC_MUTANT == 4
        const double nvx = (1. - r) * (s->vx[p] + flx) + r * picx;
        const double nvy = (1. - r) * (s->vy[p] + fly) + r * picy;
#elif ORC_MUTANT == 5
        const double nvx = r * s->vx[p] + (1. - r) * picx;
        const double nvy = r * s->vy[p] + (1. - r) * picy;
#else
        /* simulation.cpp:932 */
        const double nvx = r * (s->vx[p] + flx) + (1. - r) * picx;
        const double nvy = r * (s->vy[p] + fly) + (1. - r) * picy;
#endif
        s->vx[p] = nvx;
        s->vy[p] = nvy;
        s->vz[p] = nvx * Gx + nvy * Gy;
        s->x[p] = xp + dt * picx;
        s->y[p] = yp + dt * picy;
#if ORC_MUTANT == 3
        s->h[p] *= 1. + tr;
#else
        s->h[p] /= 1. + tr; /* simulation.cpp:937 */
#endif
        s->A[p] = s->m[p] / (s->h[p] * rho);
    }
}

void orc_prime(orc_sim* s) {
    /* simulation.cpp:493-495 */
    remesh(s);
    p2g(s);
    grid_update(s);
}

/* simulation.cpp:500-510 and advanceStep 524-563 */
int orc_step(orc_sim* s) {
    if (!(s->time < s->p.total_time)) return -1;
    s->step++;
    double t0 = now_s();
    update_dt(s);
    double t1 = now_s();
    remesh(s);
    double t2 = now_s();
    if (s->sort_enabled && s->step % 50 == 0) sort_particles(s);
    double t3 = now_s();
    p2g(s);
    double t4 = now_s();
    grid_update(s);
    double t5 = now_s();
    g2p(s);
    double t6 = now_s();
    s->t_phase[0] += t1 - t0;
    s->t_phase[1] += t2 - t1;
    s->t_phase[2] += t3 - t2;
    s->t_phase[3] += t4 - t3;
    s->t_phase[4] += t5 - t4;
    s->t_phase[5] += t6 - t5;
    s->time += s->dt;
    if (s->p.output_dt * (double)(s->frame + 1) - s->time < 1e-13) {
        s->frame++;
        return 1;
    }
    return 0;
}

int64_t orc_run(orc_sim* s, int64_t max_steps) {
    int64_t n = 0;
    while (n < max_steps && s->time < s->p.total_time) {
        orc_step(s);
        ++n;
    }
    return n;
}

void orc_scalars(const orc_sim* s, double out[10]) {
    out[0] = s->time;
    out[1] = s->dt;
    out[2] = (double)s->frame;
    out[3] = (double)s->step;
    for (int k = 0; k < 6; ++k) out[4 + k] = s->t_phase[k];
}

void orc_window(const orc_sim* s, int32_t out[4]) {
    out[0] = s->iL;
    out[1] = s->jB;
    out[2] = s->gnx;
    out[3] = s->gny;
}

static void cp(double* dst, const double* src, size_t n) {
    if (dst && src) memcpy(dst, src, n * 8);
}

void orc_get_grid(const orc_sim* s, double* mass, double* vx, double* vy, double* vz, double* h, double* flipx,
                  double* flipy, double* Fx, double* Fy) {
    size_t n = (size_t)s->gnx * s->gny;
    cp(mass, s->gm, n); cp(vx, s->gvx, n); cp(vy, s->gvy, n); cp(vz, s->gvz, n); cp(h, s->gh, n);
    cp(flipx, s->gflx, n); cp(flipy, s->gfly, n); cp(Fx, s->gFx, n); cp(Fy, s->gFy, n);
}

void orc_get_particles(const orc_sim* s, double* x, double* y, double* vx, double* vy, double* vz, double* z,
                       double* h, double* A, double* m) {
    for (int32_t k = 0; k < s->np; ++k) {
        int32_t q = s->pid[k];
        if (x) x[q] = s->x[k];
        if (y) y[q] = s->y[k];
        if (vx) vx[q] = s->vx[k];
        if (vy) vy[q] = s->vy[k];
        if (vz) vz[q] = s->vz[k];
        if (z) z[q] = s->pz[k];
        if (h) h[q] = s->h[k];
        if (A) A[q] = s->A[k];
        if (m) m[q] = s->m[k];
    }
}

void orc_set_velocity(orc_sim* s, const double* vx, const double* vy) {
    for (int32_t k = 0; k < s->np; ++k) {
        s->vx[k] = vx[s->pid[k]];
        s->vy[k] = vy[s->pid[k]];
    }
}

void orc_get_max(const orc_sim* s, double* h_max, double* v_max2) {
    size_t n = (size_t)s->nx * s->ny;
    cp(h_max, s->h_max, n);
    cp(v_max2, s->v_max2, n);
}

void orc_get_terrain(const orc_sim* s, double* gx, double* gy, double* theta) {
    size_t n = (size_t)s->nx * s->ny;
    cp(gx, s->gx, n);
    cp(gy, s->gy, n);
    cp(theta, s->theta, n);
}

void orc_set_sort(orc_sim* s, int enabled) { s->sort_enabled = enabled; }

void orc_set_time(orc_sim* s, double time, int32_t frame) {
    s->time = time;
    s->frame = frame;
}

/* ---------------- front-end restatements ---------------- */

/* simulation.cpp:478-482.  cos and sin of one angle come from one sincos()
 * call, as a GCC build of the reference computes them (GCC merges the pair
 * into sincos at -O1 and above); glibc's sincos and cos/sin can differ in
 * the last bit. */
static void rotate(double px, double py, double ang, double* ox, double* oy) {
    double c, sn;
    sincos(ang, &sn, &c);
    *ox = c * px - sn * py;
    *oy = sn * px + c * py;
}

/* polygon_tools.cpp:87-101 */
static int pip(double x, double y, const double* vx, const double* vy, int32_t n) {
    int inside = 0;
    for (int32_t i = 0, j = n - 1; i < n; j = i++) {
        const double xi = vx[i], yi = vy[i], xj = vx[j], yj = vy[j];
        int intersect = ((yi > y) != (yj > y)) && (x < (xj - xi) * (y - yi) / ((yj - yi) + 1e-12) + xi);
        if (intersect) inside = !inside;
    }
    return inside;
}

/* Samples of simulation.cpp:355-385 over bbox [xmin,xmax]x[ymin,ymax] (unsnapped);
 * keep(k) decides; order i-outer, j, then the 4 in-cell offsets. */
typedef int (*keep_fn)(const void* ctx, double x, double y);

static int64_t uniform_sample(double dx, double xmin, double xmax, double ymin, double ymax, keep_fn keep,
                              const void* ctx, double* ox, double* oy, int64_t cap) {
    const double inv = 1. / dx;
    xmin = floor(xmin * inv) * dx;
    xmax = ceil(xmax * inv) * dx;
    ymin = floor(ymin * inv) * dx;
    ymax = ceil(ymax * inv) * dx;
    int32_t ncx = (int32_t)round((xmax - xmin) * inv);
    int32_t ncy = (int32_t)round((ymax - ymin) * inv);
    static const double ddx[4] = {0.25, 0.75, 0.25, 0.75};
    static const double ddy[4] = {0.25, 0.25, 0.75, 0.75};
    int64_t n = 0;
    for (int32_t i = 0; i < ncx; ++i) {
        for (int32_t j = 0; j < ncy; ++j) {
            for (int d = 0; d < 4; ++d) {
                double px = ((double)i + ddx[d]) * dx + xmin;
                double py = ((double)j + ddy[d]) * dx + ymin;
                if (keep(ctx, px, py)) {
                    if (n < cap) {
                        ox[n] = px;
                        oy[n] = py;
                    }
                    ++n;
                }
            }
        }
    }
    return n;
}

typedef struct { int kind; double cx, cy, a, b, rot; } shape_ctx;

/* simulation.cpp:400-409 */
static int keep_shape(const void* vctx, double x, double y) {
    const shape_ctx* c = (const shape_ctx*)vctx;
    double lx, ly;
    rotate(x - c->cx, y - c->cy, -c->rot, &lx, &ly);
    if (c->kind == 0) return fabs(lx) <= c->a / 2. && fabs(ly) <= c->b / 2.;
    return pow(lx / c->a, 2.) + pow(ly / c->b, 2.) <= 1.;
}

int64_t orc_uniform_shape(double dx, int32_t kind, double cx, double cy, double a, double b, double rot_rad,
                          double* x, double* y, int64_t cap) {
    shape_ctx c = {kind, cx, cy, a, b, rot_rad};
    double xmin, xmax, ymin, ymax;
    if (kind == 0) { /* 268-275 */
        double diag = sqrt(a * a + b * b) / 2.;
        xmin = cx - diag; xmax = cx + diag; ymin = cy - diag; ymax = cy + diag;
    } else { /* 278-285 */
        double big = a > b ? a : b;
        xmin = cx - big; xmax = cx + big; ymin = cy - big; ymax = cy + big;
    }
    return uniform_sample(dx, xmin, xmax, ymin, ymax, keep_shape, &c, x, y, cap);
}

typedef struct { const double *vx, *vy; int32_t n; } poly_ctx;

static int keep_poly(const void* vctx, double x, double y) {
    const poly_ctx* c = (const poly_ctx*)vctx;
    return pip(x, y, c->vx, c->vy, c->n);
}

int64_t orc_uniform_polygon(double dx, const double* vx, const double* vy, int32_t nv, double* x, double* y,
                            int64_t cap) {
    double xmin = 1e308, ymin = 1e308, xmax = -1e308, ymax = -1e308;
    for (int32_t k = 0; k < nv; ++k) { /* 301-306 */
        if (vx[k] < xmin) xmin = vx[k];
        if (vx[k] > xmax) xmax = vx[k];
        if (vy[k] < ymin) ymin = vy[k];
        if (vy[k] > ymax) ymax = vy[k];
    }
    poly_ctx c = {vx, vy, nv};
    return uniform_sample(dx, xmin, xmax, ymin, ymax, keep_poly, &c, x, y, cap);
}

/* complex_terrain.cpp:132-183 */
double orc_raster_height(const double* z, int32_t ncols, int32_t nrows, double ox, double oy, double cs,
                         double nodata, double px, double py) {
    if (ncols == 0 || nrows == 0) return 0.0;
    double xr = px - ox, yr = py - oy;
    double xback = (double)(ncols - 1) * cs, yback = (double)(nrows - 1) * cs;
    /* x.back() - x[0] is the reference's bound only when x[0]=0; it compares
     * x_rel against x.back() itself (150), restated literally: */
    xback = ox + xback;
    yback = oy + yback;
    if (xr < 0 || yr < 0 || xr >= xback || yr >= yback) return 0.0;
    double dxl = cs, dyl = cs; /* x[1]-x[0] */
    dxl = (ox + cs) - ox;
    dyl = (oy + cs) - oy;
    int32_t i = (int32_t)floor(xr / dxl), j = (int32_t)floor(yr / dyl);
    if (i < 0 || j < 0 || i >= ncols - 1 || j >= nrows - 1) return 0.0;
    double fx = (xr - i * dxl) / dxl, fy = (yr - j * dyl) / dyl;
    double z00 = z[(size_t)i * nrows + j], z10 = z[(size_t)(i + 1) * nrows + j];
    double z01 = z[(size_t)i * nrows + j + 1], z11 = z[(size_t)(i + 1) * nrows + j + 1];
    if (fabs(z00 - nodata) < 1e-6 || fabs(z10 - nodata) < 1e-6 || fabs(z01 - nodata) < 1e-6 ||
        fabs(z11 - nodata) < 1e-6)
        return 0.0;
    double z0 = z00 * (1.0 - fx) + z10 * fx;
    double z1 = z01 * (1.0 - fx) + z11 * fx;
    double hgt = z0 * (1.0 - fy) + z1 * fy;
    return hgt > 0.0 ? hgt : 0.0;
}
