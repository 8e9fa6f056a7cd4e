/*
 * oracle.h -- C API of the CPU oracle for the Smilei PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in plain
 * FP64, the arithmetic and the per-patch structure of Smilei's time step
 * (Species::dynamics -> Interpolator -> PusherBoris -> Projector, particle
 * exchange + cell sort, J ghost sum, Yee Maxwell-Ampere / Faraday, B
 * exchange / centring).  It is the checker for the HIP library and the CPU
 * baseline of bench.py; no product code may call it.
 *
 * Field ids used by ora_set_field / ora_get_field (global periodic arrays,
 * row-major [Nx][Ny][Nz], Nz = 1 in 2D; a dual index g sits at (g-0.5)*dx):
 *   0 Ex, 1 Ey, 2 Ez, 3 Bx, 4 By, 5 Bz, 6 Bx_m, 7 By_m, 8 Bz_m,
 *   9 Jx, 10 Jy, 11 Jz
 */
#ifndef SMILEI_ORACLE_H
#define SMILEI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    ndim;          /* 2 or 3                                        */
    int    order;         /* shape order of interpolator/projector: 2 or 4 */
    int    n_cells[3];    /* global cells per dimension (1 in unused dims) */
    int    patch_size[3]; /* cells per patch                               */
    double cell_length[3];
    double timestep;
} ora_params;

/* ---- whole-simulation API (per-patch reference structure) ------------ */
void   *ora_create(const ora_params *p);
void    ora_destroy(void *h);
void    ora_set_threads(int n);
int     ora_add_species(void *h, double mass, double charge);
/* Species::time_frozen_: no push, deposit or exchange while the step's
 * time_dual = (steps done + 1.5) * timestep <= time_frozen. */
int     ora_set_time_frozen(void *h, int sp, double time_frozen);
int     ora_set_particles(void *h, int sp, int64_t n,
                          const double *x, const double *y, const double *z,
                          const double *px, const double *py, const double *pz,
                          const double *w);
int64_t ora_num_particles(void *h, int sp);
/* Particles in canonical order: patch-major (patch id = (px*npy+py)*npz+pz),
 * then the patch's stored order (sorted by cell key).  Any pointer may be
 * NULL. patch/key receive the owning patch and the in-patch cell key. */
int     ora_get_particles(void *h, int sp, double *x, double *y, double *z,
                          double *px, double *py, double *pz, double *w,
                          int32_t *patch, int32_t *key);
int     ora_patch_counts(void *h, int sp, int32_t *counts);
int     ora_set_field(void *h, int fid, const double *global);
int     ora_get_field(void *h, int fid, double *global);
int     ora_step(void *h, int nsteps);
/* out[0] = Uelm, out[1+s] = Ukin of species s */
int     ora_energies(void *h, double *out);
/* out[0..5] = Uelm of Ex, Ey, Ez, Bx_m, By_m, Bz_m: 0.5 * V_cell * the
 * Field3D::norm2 sums over each patch's istart/bufsize ranges. */
int     ora_field_energies(void *h, double *out);
/* Charge density of all species (order-matched shape, reference
 * Projector*::basic type 0), ghost-summed, as a global primal array. */
int     ora_rho(void *h, double *global);
/* Gauss residual field div(E) - rho on primal nodes (global array). */
int     ora_gauss(void *h, double *global);
/* Timers of the last ora_step call, seconds: [0] particles (gather, push,
 * deposit), [1] exchange+sort, [2] J ghost sum, [3] Maxwell + B sync. */
int     ora_timers(void *h, double *out);

/* ---- multi-patch synchronisation on patch-major flat arrays ----------- *
 * Patch p's block starts at p * (|X| + |Y| + |Z|) and holds the three
 * components in the per-patch ghosted layout (patch ids (px*npy+py)*npz+pz). */
/* SyncVectorPatch::sumAllComponents of Jx, Jy, Jz (x, y, z passes). */
int ora_sum_j_ghosts(const ora_params *p, double *J);
/* SyncVectorPatch::exchangeB of Bx, By, Bz (x, y, z passes). */
int ora_exchange_b(const ora_params *p, double *B);

/* ---- single-patch operators (reference per-patch layout) -------------- *
 * Field arrays are the patch's ghosted arrays (primal dims np = patch +
 * 2*order + 1, dual = np + 1 along the staggered axis), dom_begin is the
 * patch's cell_starting_global_index (Pcoord*patch - oversize).          */
void ora_interp3d(int order, const double *Ex, const double *Ey, const double *Ez,
                  const double *Bxm, const double *Bym, const double *Bzm,
                  const int np[3], const int dom_begin[3], const double dinv[3],
                  int64_t n, const double *x, const double *y, const double *z,
                  double *Epart, double *Bpart, int32_t *iold, double *deltaold);
void ora_interp2d(int order, const double *Ex, const double *Ey, const double *Ez,
                  const double *Bxm, const double *Bym, const double *Bzm,
                  const int np[2], const int dom_begin[2], const double dinv[2],
                  int64_t n, const double *x, const double *y,
                  double *Epart, double *Bpart, int32_t *iold, double *deltaold);
void ora_push_boris(int ndim, double charge, double mass, double dt, int64_t n,
                    double *x, double *y, double *z,
                    double *px, double *py, double *pz,
                    const double *Epart, const double *Bpart, double *invgf);
void ora_proj3d(int order, double *Jx, double *Jy, double *Jz,
                const int np[3], const int dom_begin[3], const double dx[3],
                double dt, double inv_cell_volume, double charge, int64_t n,
                const double *x, const double *y, const double *z,
                const double *w, const int32_t *iold, const double *deltaold);
void ora_proj2d(int order, double *Jx, double *Jy, double *Jz,
                const int np[2], const int dom_begin[2], const double dx[2],
                double dt, double inv_cell_volume, double charge, int64_t n,
                const double *x, const double *y, const double *pz,
                const double *w, const double *invgf,
                const int32_t *iold, const double *deltaold);
void ora_rho3d(int order, double *rho, const int np[3], const int dom_begin[3],
               const double dinv[3], double inv_cell_volume, double charge,
               int64_t n, const double *x, const double *y, const double *z,
               const double *w);
void ora_rho2d(int order, double *rho, const int np[2], const int dom_begin[2],
               const double dinv[2], double inv_cell_volume, double charge,
               int64_t n, const double *x, const double *y, const double *w);
void ora_yee3d_ma(double *Ex, double *Ey, double *Ez,
                  const double *Bx, const double *By, const double *Bz,
                  const double *Jx, const double *Jy, const double *Jz,
                  const int np[3], double dt, const double dx[3]);
void ora_yee3d_mf(double *Bx, double *By, double *Bz,
                  const double *Ex, const double *Ey, const double *Ez,
                  const int np[3], double dt, const double dx[3]);
void ora_yee2d_ma(double *Ex, double *Ey, double *Ez,
                  const double *Bx, const double *By, const double *Bz,
                  const double *Jx, const double *Jy, const double *Jz,
                  const int np[2], double dt, const double dx[2]);
void ora_yee2d_mf(double *Bx, double *By, double *Bz,
                  const double *Ex, const double *Ey, const double *Ez,
                  const int np[2], double dt, const double dx[2]);

#ifdef __cplusplus
}
#endif
#endif
