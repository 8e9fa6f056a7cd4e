/*
 * smilei_amd.h -- C ABI of the MI355X-native PIC hot path.
 *
 * The library replaces Smilei's per-timestep particle loop
 *   Species::dynamics -> Interpolator::fieldsWrapper -> Pusher::operator()
 *                     -> Projector::currentsAndDensityWrapper
 * plus the per-patch cell sort / exchange and the Yee field solver
 * (MA_Solver3D_norm, MF_Solver3D_Yee, their 2D forms), with hand-written HIP
 * kernels for gfx950.  Two layers are exported:
 *
 *  1. A context ("whole time step") API: particles and fields live in HBM,
 *     smi_step() runs N complete PIC steps, batching every patch of every
 *     species into single launches (optionally one captured hipGraph per
 *     step).  This is the engine behind VectorPatch::dynamics /
 *     sumDensities / solveMaxwell / finalizeExchParticlesAndSort
 *     (src/Patch/VectorPatch.cpp:325, 819, 965; src/Patch/SyncVectorPatch.cpp:44).
 *
 *  2. Per-operator entry points on one patch in the reference's own memory
 *     layout (ghosted per-patch Field3D/Field2D arrays, SoA particles,
 *     SmileiMPI dynamics_* scratch), for a reference build that keeps its
 *     own orchestration and swaps the operators only.
 *
 * All functions return SMI_OK (0) or a negative error code; smi_last_error()
 * returns a text for the last failure of the calling thread.  Pointers named
 * d_* are device pointers; all other array pointers are host pointers.
 * Calls on one context must come from one host thread (the reference's GPU
 * build uses one OpenMP thread per rank as well, src/Smilei.cpp:115-134).
 */
#ifndef SMILEI_AMD_H
#define SMILEI_AMD_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define SMI_OK 0
#define SMI_ERR_INVALID (-1)  /* bad argument or unsupported namelist value */
#define SMI_ERR_HIP (-2)      /* HIP runtime failure                        */
#define SMI_ERR_STATE (-3)    /* call not valid in the current state         */
#define SMI_ERR_CAPACITY (-4) /* a size exceeds what was allocated           */

/* Precision forms of the particle path. */
#define SMI_PLAIN_FP64 0 /* FP64 everywhere (the drop-in form)                 */
#define SMI_MIXED 1      /* FP32 particle arithmetic, int32 cell + FP32 offset
                            positions, FP64 fields and FP64 J reduction       */

/* Current deposition modes (plain FP64 form). */
#define SMI_DEPOSIT_ATOMIC 0        /* LDS tile per 4^3 cluster + FP64 atomics */
#define SMI_DEPOSIT_DETERMINISTIC 1 /* per-patch ghosted J, particle order and
                                       ghost-sum order of the reference        */

/* Field identifiers (global periodic arrays [Nx][Ny][Nz], Nz=1 in 2D; a dual
 * index g along an axis sits at (g - 1/2) * cell_length, like Smilei's). */
#define SMI_EX 0
#define SMI_EY 1
#define SMI_EZ 2
#define SMI_BX 3
#define SMI_BY 4
#define SMI_BZ 5
#define SMI_BXM 6
#define SMI_BYM 7
#define SMI_BZM 8
#define SMI_JX 9
#define SMI_JY 10
#define SMI_JZ 11
#define SMI_NFIELDS 12

typedef struct smi_context smi_context;

/* Namelist subset consumed by the factories.  The string fields take the
 * reference's namelist values and are checked the way its factories check
 * them (InterpolatorFactory.h:44, PusherFactory.h:41, ProjectorFactory.h:32,
 * SolverFactory.h:55,128); NULL selects the reference default. */
typedef struct {
    const char *geometry;       /* "2Dcartesian" | "3Dcartesian"           */
    int interpolation_order;    /* 2 | 4 (4: 3Dcartesian only)             */
    const char *interpolator;   /* "momentum-conserving" (default)         */
    const char *pusher;         /* "boris" (default)                        */
    const char *maxwell_solver; /* "Yee" (default)                          */
    int n_cells[3];             /* global cells (Main.grid_length / cell_length) */
    int patch_size[3];          /* cells per patch (multiple of 4)          */
    double cell_length[3];
    double timestep;
    int precision;              /* SMI_PLAIN_FP64 | SMI_MIXED               */
    int deposit;                /* SMI_DEPOSIT_ATOMIC | _DETERMINISTIC       */
    int use_graph;              /* capture each time step in a hipGraph     */
    int device;                 /* HIP device ordinal                      */
} smi_params;

const char *smi_last_error(void);
/* Hash of the sources the library was built from (16 hex digits); no
 * counterpart in the reference, whose build has none.  The Python loader
 * refuses a library whose hash differs from the sources beside it. */
const char *smi_build_hash(void);

/* ---- context API --------------------------------------------------- */
int smi_create(const smi_params *params, smi_context **out);
int smi_destroy(smi_context *ctx);
/* mass in m_e, charge in e (integer-valued like Particles::charge);
 * capacity = maximum number of macro-particles of the species. */
int smi_add_species(smi_context *ctx, const char *name, double mass, double charge,
                    int64_t capacity, int *species_id);
/* Load particles (global coordinates, momenta u = gamma*beta in units of
 * m_species*c, weights in the reference convention n*V_cell/ppc).
 * SMI_ERR_INVALID, with the species left as it was, for a position outside
 * [0, grid length) or a non-finite momentum or weight. */
int smi_set_particles(smi_context *ctx, int species, int64_t n,
                      const double *x, const double *y, const double *z,
                      const double *px, const double *py, const double *pz,
                      const double *w);
/* Seeded in-HBM initialisation of a species: `ppc` particles per cell,
 * uniform random in each cell, Gaussian momenta of std `sigma` (units of
 * m_species*c) per component, weight `density*V/ppc`.  A drift u_d =
 * drift_ux = gamma_d*v_d along x Lorentz-boosts the drawn momenta as the
 * reference's boostParticles does (ParticleCreator.cpp:755-801: Zenitani's
 * volume-transform correction, then u_x' = g u_x + gamma u_d); cold
 * particles get u_x = u_d exactly.
 * Cells whose centre lies outside the cylinder of radius `radius` (in
 * cells, around the (y,z) axis centre; <= 0 disables) get no particles. */
int smi_init_species(smi_context *ctx, int species, int ppc, double density,
                     double sigma, double drift_ux, double radius, uint64_t seed);
/* Same with isotropic Maxwell-Juttner momenta of temperature theta =
 * T / (m_species c^2) (momentum_initialization = "maxwell-juttner" of the
 * reference, ParticleCreator.cpp:1002, in the rest frame, then boosted by
 * drift_ux as above), drawn by inverse transform from 4096 quantiles of |u|
 * tabulated on the host; theta = 0 gives cold particles. */
int smi_init_species_mj(smi_context *ctx, int species, int ppc, double density,
                        double theta, double drift_ux, double radius, uint64_t seed);
/* Particles at the positions of species `src` (weights times weight_scale),
 * fresh Maxwell-Juttner momenta of temperature theta (0: cold) boosted by
 * `drift_ux` along x: position_initialization = "<species name>" of the
 * reference (src/Particles/ParticleCreator.cpp:277-284). */
int smi_init_species_like(smi_context *ctx, int species, int src, double weight_scale, double theta,
                          double drift_ux, uint64_t seed);
/* Profile-driven initialisation, ParticleCreator::create
 * (src/Particles/ParticleCreator.cpp:151-346): host arrays [ncells] of the
 * number density and particles-per-cell profiles evaluated at the cell
 * centres (Profile::valuesAt); a fractional ppc adds one particle on a
 * regular lattice of cells (:217-231), cells with zero density or ppc get
 * none (:242-246), weight = density * V_cell / nPart (createWeight :936).
 * Positions uniform in the cell, Gaussian momenta of std `sigma` plus
 * `drift` along x. */
int smi_init_species_profile(smi_context *ctx, int species, const double *density, const double *ppc,
                             double sigma, double drift_ux, uint64_t seed);
/* Same with Maxwell-Juttner momenta of temperature theta (as smi_init_species_mj). */
int smi_init_species_profile_mj(smi_context *ctx, int species, const double *density, const double *ppc,
                                double theta, double drift_ux, uint64_t seed);
int64_t smi_num_particles(const smi_context *ctx, int species);
/* Particles in canonical order (patch-major, then cell key, then arrival
 * order; in the mixed form, whose storage is binned by the nearest grid
 * node, the particles of a cell come by position, x then y then z); patch
 * and key receive the owning patch id ((px*npy+py)*npz+pz) and the in-patch
 * cell key of the cell the position lies in.  Any pointer may be NULL. */
int smi_get_particles(smi_context *ctx, int species, double *x, double *y, double *z,
                      double *px, double *py, double *pz, double *w,
                      int32_t *patch, int32_t *key);
int smi_patch_counts(smi_context *ctx, int species, int32_t *counts);
int smi_set_field(smi_context *ctx, int field, const double *global);
int smi_get_field(smi_context *ctx, int field, double *global);
/* Run n complete time steps (stream-ordered; returns after enqueueing). */
int smi_step(smi_context *ctx, int nsteps);
int smi_synchronize(smi_context *ctx);
/* Species::time_frozen_ (src/Species/Species.cpp:568, 832-866): the
 * species is not pushed, deposits no current and is not re-sorted while
 * the step's time_dual = (steps done + 1.5) * timestep (Smilei.cpp:172,
 * 487-488) is <= time_frozen; it still counts in smi_rho / smi_gauss and
 * the energies.  0 (the default) never freezes. */
int smi_set_time_frozen(smi_context *ctx, int species, double time_frozen);
/* Steps done so far (smi_step advances it); a restart sets it so that
 * time-dependent choices (frozen species) continue where the dump left. */
int smi_set_step_count(smi_context *ctx, int64_t steps_done);
int64_t smi_step_count(const smi_context *ctx);
/* out[0] = Uelm, out[1+s] = Ukin of species s (DiagnosticScalar formulas). */
int smi_energies(smi_context *ctx, double *out);
/* The per-patch sums of DiagnosticScalar::compute
 * (src/Diagnostic/DiagnosticScalar.cpp:522-660) over the whole domain:
 * out[0..5] = Uelm_Ex, Uelm_Ey, Uelm_Ez, Uelm_Bx_m, Uelm_By_m, Uelm_Bz_m
 * (0.5 * V_cell * sum of squares, Field3D::norm2), then for each species s
 * out[6+3s..] = Ntot (particles), Dens (sum of weights), Ukin (mass * sum
 * w (gamma - 1)).  The energy balance (Utot, Uexp, Ubal) is formed by the
 * caller from these, as SmileiMPI::computeGlobalDiags does. */
int smi_scalars(smi_context *ctx, double *out);
/* Charge density (all species) and div(E) - rho on primal nodes. */
int smi_rho(smi_context *ctx, double *global);
int smi_gauss(smi_context *ctx, double *global);
/* Per-kernel-class device time of one step, measured with HIP events on the
 * context's stream: ms[0] particle kernel (gather+push+deposit), ms[1] sort,
 * ms[2] J reduction (deterministic mode), ms[3] Maxwell. */
int smi_profile_step(smi_context *ctx, float *ms);
/* The stream all work of the context is enqueued on (hipStream_t). */
void *smi_stream(smi_context *ctx);
/* Launch-shape knobs of the particle kernels, searched by smi_autotune (no
 * effect on results beyond the order of the atomic J summation).
 * SMI_TUNE_INTERLEAVE (mixed form and atomic FP64 form): 1 = the sort
 * stores each cluster's particles lane-interleaved (64 columns of
 * consecutive cell-sorted ranks, stored row by row), so that the 64 lanes of
 * a wavefront load consecutive slots and hold particles of distinct cells;
 * 0 = cell-major storage (lanes share cells); 2 (mixed form) = pair-interleaved (two
 * consecutive ranks in the adjacent lanes 2c, 2c+1, the 32 pairs of a row in
 * distinct cells): the mixed kernel then adds the J-tile words of the two
 * lanes of a pair that share a node across the lanes (DPP) and issues one
 * LDS atomic per pair.  Changing it re-sorts the current state;
 * smi_get_particles returns cell-major order either way.
 * SMI_TUNE_THREADS: workgroup size per 4^3
 * cluster, 64 | 128 | 256 | 512 (particles per thread = cluster particles /
 * threads).  SMI_TUNE_PERSISTENT (both forms): 0 = one workgroup per
 * cluster of the grid, 1 = a grid of the resident workgroups pulling the
 * non-empty clusters from a work list built each step. */
#define SMI_TUNE_INTERLEAVE 1
#define SMI_TUNE_THREADS 2
#define SMI_TUNE_PERSISTENT 3
/* SMI_TUNE_XCD: 1 = the grid launch hands each of the chip's 8 XCDs (own L2
 * each; workgroups are dispatched to them round-robin) a contiguous run of
 * clusters, so that neighbouring clusters share L2 lines; 0 = grid order. */
#define SMI_TUNE_XCD 5
/* SMI_TUNE_TILE (mixed form): LDS tile shape of the particle kernel, 1 = J
 * rows padded by one 64-bit word (fewer bank conflicts of the tile atomics
 * on dense plasmas), 0 = unpadded. */
#define SMI_TUNE_TILE 6
/* SMI_TUNE_TILE_X (mixed form, order 2): the patch tile one workgroup
 * runs, in 4^3 clusters along x: 1 (default) or 2 (an 8x4x4-cell tile: one
 * field tile and one J tile for two clusters, fewer halo points loaded and
 * flushed per particle, two 512-thread workgroups per CU).  2 needs patches
 * of a multiple of 8 cells along x; the persistent form runs one cluster
 * per workgroup whatever this knob says. */
#define SMI_TUNE_TILE_X 7
/* SMI_TUNE_NODE_ACC (mixed form, order 2): 1 = the particles that keep their
 * node are staged per node in LDS, 256 at a time, and their J-tile words
 * summed node by node in registers, one LDS atomic per node and word (three
 * workgroups per CU for the staging); 0 (default) = one atomic per particle
 * and word.  The same fixed-point words either way.  Runs with 256-thread
 * one-cluster grid launches and the lane-interleaved or cell-major layouts;
 * other launch shapes deposit by particle whatever this knob says. */
#define SMI_TUNE_NODE_ACC 8
/* Not searched by the autotuner (it changes what smi_get_particles returns):
 * SMI_TUNE_CELL_ORDER 1 (default) keeps the canonical order of particles
 * inside a cell (direction of motion, then previous slot); 0 leaves it to
 * the atomic scatter, like the reference's GPU sort (thrust sort_by_key,
 * nvidiaParticles.cu:1217-1247), and saves the in-cell sort pass.  The
 * deterministic deposit always keeps the canonical order. */
#define SMI_TUNE_CELL_ORDER 4
int smi_set_tuning(smi_context *ctx, int key, int value);
int smi_get_tuning(smi_context *ctx, int key, int *value);
/* Project autotuner (no counterpart in the reference, whose GPU launch shapes
 * are fixed): times the particle-kernel launches of one step on the live
 * particle state for every lane mapping x launch form (mixed form) x
 * deposit by particle or node (SMI_TUNE_NODE_ACC) x
 * patch-tile extent (SMI_TUNE_TILE_X) x cluster order (SMI_TUNE_XCD) x
 * workgroup size 128/256/512, `reps`
 * launches each, and keeps the fastest; best_ms (may be
 * NULL) receives its time.  Leaves the simulation state as it was. */
int smi_autotune(smi_context *ctx, int reps, float *best_ms);

/* ---- per-operator entry points (one patch, reference layout) -------- *
 * Field arrays are the patch's ghosted arrays: primal extent
 * np[d] = patch_size[d] + 2*oversize + 1, +1 along each staggered axis
 * (ElectroMagn3D.cpp:88-180); dom_begin = cell_starting_global_index
 * (Patch.cpp:145-162).  Scratch arrays have stride n as SmileiMPI's
 * dynamics_Epart/Bpart/iold/deltaold/invgf (SmileiMPI.h:213-221).      */

/* Interpolator3D2Order / Interpolator3D4Order ::fieldsWrapper
 * (src/Interpolator/Interpolator3D2Order.cpp:163-285,
 *  src/Interpolator/Interpolator3D4Order.cpp:159-). */
int smi_interpolate_3d(int order, const double *d_Ex, const double *d_Ey, const double *d_Ez,
                       const double *d_Bxm, const double *d_Bym, const double *d_Bzm,
                       const int np[3], const int dom_begin[3], const double cell_length[3],
                       int64_t n, const double *d_x, const double *d_y, const double *d_z,
                       double *d_Epart, double *d_Bpart, int32_t *d_iold, double *d_deltaold,
                       void *stream);
/* Interpolator2D2Order::fieldsWrapper (src/Interpolator/Interpolator2D2Order.cpp:158-). */
int smi_interpolate_2d(int order, const double *d_Ex, const double *d_Ey, const double *d_Ez,
                       const double *d_Bxm, const double *d_Bym, const double *d_Bzm,
                       const int np[2], const int dom_begin[2], const double cell_length[2],
                       int64_t n, const double *d_x, const double *d_y,
                       double *d_Epart, double *d_Bpart, int32_t *d_iold, double *d_deltaold,
                       void *stream);
/* PusherBoris::operator() (src/Pusher/PusherBoris.cpp:23-142). */
int smi_push_boris(int ndim, double charge, double mass, double timestep, int64_t n,
                   double *d_x, double *d_y, double *d_z,
                   double *d_px, double *d_py, double *d_pz,
                   const double *d_Epart, const double *d_Bpart, double *d_invgf, void *stream);
/* Projector3D2Order / Projector3D4Order ::currents over particles [0,n)
 * (src/Projector/Projector3D2Order.cpp:55-343, 733-766;
 *  src/Projector/Projector3D4Order.cpp:49-237, 682-716); replaces the
 * reference's extern "C" currentDeposition3DOnDevice
 * (src/Projector/Projector3D2OrderGPU.cpp:55-88).  deterministic != 0 adds
 * the particles in index order (bit-identical to the CPU operator),
 * otherwise FP64 atomics. */
int smi_currents_3d(int order, double *d_Jx, double *d_Jy, double *d_Jz,
                    const int np[3], const int dom_begin[3], const double cell_length[3],
                    double timestep, double charge, int64_t n,
                    const double *d_x, const double *d_y, const double *d_z, const double *d_w,
                    const int32_t *d_iold, const double *d_deltaold, int deterministic,
                    void *stream);
/* Projector2D2Order::currents (src/Projector/Projector2D2Order.cpp:49-189). */
int smi_currents_2d(int order, double *d_Jx, double *d_Jy, double *d_Jz,
                    const int np[2], const int dom_begin[2], const double cell_length[2],
                    double timestep, double charge, int64_t n,
                    const double *d_x, const double *d_y, const double *d_pz, const double *d_w,
                    const double *d_invgf, const int32_t *d_iold, const double *d_deltaold,
                    int deterministic, void *stream);
/* MA_Solver3D_norm / MF_Solver3D_Yee ::operator()
 * (src/ElectroMagnSolver/MA_Solver3D_norm.cpp:18-115,
 *  src/ElectroMagnSolver/MF_Solver3D_Yee.cpp:17-110). */
int smi_maxwell_ampere_3d(double *d_Ex, double *d_Ey, double *d_Ez,
                          const double *d_Bx, const double *d_By, const double *d_Bz,
                          const double *d_Jx, const double *d_Jy, const double *d_Jz,
                          const int np[3], double timestep, const double cell_length[3], void *stream);
int smi_maxwell_faraday_3d(double *d_Bx, double *d_By, double *d_Bz,
                           const double *d_Ex, const double *d_Ey, const double *d_Ez,
                           const int np[3], double timestep, const double cell_length[3], void *stream);

/* Projector3D2Order / Projector3D4Order ::basic, type 0 (charge density at
 * the particles' positions: diagnostic steps of currentsAndDensityWrapper
 * with diag_flag, frozen species; Projector3D2Order.cpp:526-621) over
 * particles [0,n) of one species, into the patch's primal rho array (or
 * rho_s of the species).  FP64 atomics: equal to the CPU operator up to the
 * order of the sums. */
int smi_density_3d(int order, double *d_rho, const int np[3], const int dom_begin[3], const double cell_length[3],
                   double charge, int64_t n, const double *d_x, const double *d_y, const double *d_z,
                   const double *d_w, void *stream);
/* Projector2D2Order::basic, type 0 (Projector2D2Order.cpp). */
int smi_density_2d(int order, double *d_rho, const int np[2], const int dom_begin[2], const double cell_length[2],
                   double charge, int64_t n, const double *d_x, const double *d_y, const double *d_w, void *stream);
/* MA_Solver2D_norm / MF_Solver2D_Yee ::operator()
 * (src/ElectroMagnSolver/MA_Solver2D_norm.cpp:18-105,
 *  src/ElectroMagnSolver/MF_Solver2D_Yee.cpp:18-99) on one patch. */
int smi_maxwell_ampere_2d(double *d_Ex, double *d_Ey, double *d_Ez,
                          const double *d_Bx, const double *d_By, const double *d_Bz,
                          const double *d_Jx, const double *d_Jy, const double *d_Jz,
                          const int np[2], double timestep, const double cell_length[2], void *stream);
int smi_maxwell_faraday_2d(double *d_Bx, double *d_By, double *d_Bz,
                           const double *d_Ex, const double *d_Ey, const double *d_Ez,
                           const int np[2], double timestep, const double cell_length[2], void *stream);

/* ---- the reference's own C seams, same argument lists ---------------- *
 * currentDeposition3DOnDevice / densityDeposition3DOnDevice
 * (src/Projector/Projector3D2OrderGPU.cpp:55-120, called from
 * Projector3D2OrderGPU::currentsAndDensityWrapper :380-487).  Every pointer
 * is a device pointer (the reference's launcher maps its host pointers with
 * HostDeviceMemoryManagement::GetDevicePointer before its kernel launch; a
 * binding passes the mapped pointers here).  The particles of one patch are
 * binned by 4^3 cluster in the x-major cluster order: bin_index[b] is the
 * end of bin b (Particles::last_index), x/y/z_dimension_bin_count the
 * clusters per axis, iold / deltaold the SmileiMPI scratch of the
 * interpolator with stride number_of_particles, not_spectral = 1 for the
 * staggered Jy / Jz of the Yee solver, cell_sorting selects the loop order
 * of a bin (results differ only by the order of the atomic sums).  Order-2
 * shape, as the reference's GPU kernels.  The work runs on the null stream
 * and has finished when the call returns, as in the reference (which
 * synchronises the device after its kernel).  Returns SMI_ERR_INVALID, and
 * skips the particle, when a particle's stencil leaves its bin's tile. */
int smi_current_deposition_3d(double *Jx, double *Jy, double *Jz, int Jx_size, int Jy_size, int Jz_size,
                              const double *particle_position_x, const double *particle_position_y,
                              const double *particle_position_z, const short *particle_charge,
                              const double *particle_weight, const int *bin_index,
                              unsigned int x_dimension_bin_count, unsigned int y_dimension_bin_count,
                              unsigned int z_dimension_bin_count, const double *invgf, const int *iold,
                              const double *deltaold, unsigned int number_of_particles, double inv_cell_volume,
                              double dx_inv, double dy_inv, double dz_inv, double dx_ov_dt, double dy_ov_dt,
                              double dz_ov_dt, int i_domain_begin, int j_domain_begin, int k_domain_begin,
                              int nprimy, int nprimz, int not_spectral, bool cell_sorting);
int smi_density_deposition_3d(double *rho, int rho_size, const double *particle_position_x,
                              const double *particle_position_y, const double *particle_position_z,
                              const short *particle_charge, const double *particle_weight, const int *bin_index,
                              unsigned int x_dimension_bin_count, unsigned int y_dimension_bin_count,
                              unsigned int z_dimension_bin_count, const double *invgf, const int *iold,
                              const double *deltaold, unsigned int number_of_particles, double inv_cell_volume,
                              double dx_inv, double dy_inv, double dz_inv, double dx_ov_dt, double dy_ov_dt,
                              double dz_ov_dt, int i_domain_begin, int j_domain_begin, int k_domain_begin,
                              int nprimy, int nprimz, int not_spectral, bool cell_sorting);

/* ---- synchronisation between the patches of one device -------------- *
 * The patches' ghosted arrays are held in one patch-major array: patch p
 * (p = (px*npy+py)*npz+pz) owns the block [p*B, (p+1)*B) holding its three
 * components one after the other, each in the per-patch layout
 * (np = patch_size + 2*order + 1 points per axis, +1 along a dual axis).
 * The domain is periodic. */
/* SyncVectorPatch::sumAllComponents for Jx, Jy, Jz (src/Patch/
 * SyncVectorPatch.cpp:203-~630): along x, then y, then z, the overlapping
 * 2*order+1(+1 dual) planes of neighbours are summed and copied back. */
int smi_sum_j_ghosts(int ndim, int order, const int n_cells[3], const int patch_size[3], double *d_J,
                     void *stream);
/* SyncVectorPatch::exchangeB for Bx, By, Bz (SyncVectorPatch.cpp:667,
 * 1441-1830): per axis, the ghost planes of the components dual along it
 * are copied from the neighbours' interiors. */
int smi_exchange_b(int ndim, int order, const int n_cells[3], const int patch_size[3], double *d_B,
                   void *stream);
/* Particle exchange between the patches and cell sort of one species on one
 * device: Patch::exchParticles / cornersParticles / importAndSortParticles
 * (src/Patch/Patch.cpp:535-819) and Species::sortParticles
 * (src/Species/Species.cpp:1046).  Input: the species' n particles of all
 * patches in any order, positions at most one box length outside the
 * periodic box (wrapped as Patch::prepareParticles does).  Output, in
 * place: the particles ordered by owning patch, then by in-patch cell key
 * (Cluster3D::CellKey, src/Particles/nvidiaParticles.cu:639-660), stable
 * inside a cell; d_key (may be NULL) the in-patch cell keys, d_patch_count
 * (may be NULL) the particles per patch, d_bin_index (may be NULL) per patch
 * the cumulative ends of its 4^3 (2D: 4^2) clusters relative to the patch's
 * first particle (Particles::last_index).  charge may be NULL; z may be
 * NULL in 2D.  Allocates its scratch; returns once the stream drained. */
int smi_sort_exchange(int ndim, const int n_cells[3], const int patch_size[3], const double cell_length[3],
                      int64_t n, double *d_x, double *d_y, double *d_z, double *d_px, double *d_py, double *d_pz,
                      double *d_w, int16_t *d_charge, int32_t *d_key, int32_t *d_patch_count,
                      int32_t *d_bin_index, void *stream);
/* No counterpart in the reference.  smi_sort_exchange without any device
 * allocation, for callers that cannot allocate (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc): the scratch lives
 * in a static 64 MiB device array of the library, one per GPU, held until
 * the process ends.  Same arguments, results and errors as
 * smi_sort_exchange; returns SMI_ERR_CAPACITY when the call needs more
 * scratch words than smi_sort_exchange_noalloc_max_words() (about
 * 2.25 words per cell of the grid plus 5.1 per particle).  Calls that
 * overlap on one GPU share the array and race. */
int smi_sort_exchange_noalloc(int ndim, const int n_cells[3], const int patch_size[3], const double cell_length[3],
                              int64_t n, double *d_x, double *d_y, double *d_z, double *d_px, double *d_py,
                              double *d_pz, double *d_w, int16_t *d_charge, int32_t *d_key, int32_t *d_patch_count,
                              int32_t *d_bin_index, void *stream);
int64_t smi_sort_exchange_noalloc_max_words(void);

#ifdef __cplusplus
}
#endif
#endif
