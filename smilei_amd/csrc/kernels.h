// kernels.h -- host-side launchers of the HIP kernels (internal C++ API).
#pragma once

#include "common.h"

namespace smi {

// Records the text returned by smi_last_error() for the calling thread.
void set_error(const std::string &msg);

struct PartPtrs {
    double *x, *y, *z, *px, *py, *pz, *w;
};

// Mixed-precision particle storage: FP32 offsets inside the cell (the cell
// itself is the key), FP32 momenta and weight.
struct MixPtrs {
    float *ox, *oy, *oz, *ux, *uy, *uz, *w;
};

// Species constants of the mixed kernel (FP32).
struct SpeciesKM {
    float qmh;          // charge/mass * dt/2, rounded to FP32
    float qmh_lo;       // qmh_fp64 - qmh (the kick applies both)
    float cw_scale;     // charge / V_cell, rounded to FP32
    float dxdt[3];      // dx/dt, dy/dt, dz/dt
    double cwd;         // charge / V_cell in FP64 (current and density weights)
    double cr_max;      // largest current contribution |q| w_max / V * max(dx/dt)
};

// Fields read by the gather (E and the centred B_m).
struct GatherFields {
    const double *Ex, *Ey, *Ez, *Bx, *By, *Bz;
};

struct JPtrs {
    double *Jx, *Jy, *Jz;
};

// ---- particle step (particles.hip) ------------------------------------
constexpr int kMaxMerged = 4;  // species per merged particle launch
// One species of a particle launch of the FP64 form (device array).
struct FpSpecies {
    SpeciesK k;
    PartPtrs A, B;                   // input / output buffers
    const int32_t *perm, *cell_off;  // sorted view of the input
    uint32_t *keys;                  // destination keys written for B
    int32_t *counts;                 // destination histogram
    double *scratch;                 // 7*cap doubles (deterministic deposit)
    int32_t *rank;                   // or null: cell-major input position of the output slots that left their cluster
    uint8_t *stay;                   // or null: placed stayers per cell (SortBufs::stay)
    int lay;                         // cluster layout of the input (layout_slot)
    int64_t cap;
};
// One block per 4-cell cluster of every patch runs the particles of the ns
// <= kMaxMerged species.  Reads each sorted particle list through `perm`
// from A, writes the advanced particles to B in sorted slot order, their
// packed destination keys (dir<<27 | global cell) and, when `deposit`, the
// Esirkepov currents into J (one LDS tile per cluster + FP64 atomics).
// When !deposit, `scratch` receives the old and the new (unwrapped)
// positions and 1/gamma for the deterministic deposit.  A particle found
// outside its cluster's stencil tile (corrupt input) is left unchanged and
// raises *err.  The kernel also accumulates the histogram of destination
// cells into `counts` (zero on entry) for launch_sort.
// persistent_blocks > 0 selects the persistent form: that many workgroups
// pull the non-empty clusters from a work list (work: one int per cluster,
// wcount: 2 ints), as for the mixed form.
void launch_particles(const Geom &g, const FpSpecies *specs, int ns, const GatherFields &F, const JPtrs &J,
                      bool deposit, int *err, int threads, bool xcd, int persistent_blocks, int *work, int *wcount,
                      hipStream_t st);
// Resident workgroups of the persistent FP64 kernel on the current device.
int fp_persistent_blocks(const Geom &g, bool deposit, int threads);
// Largest workgroup size the particle kernel of this form accepts.
int particle_max_threads(const Geom &g, bool mixed);

// Deterministic deposit into per-patch ghosted J arrays (reference layout,
// particle order of the patch).  `pj` holds npatches * (|Jx|+|Jy|+|Jz|).
struct DetSpecies {
    SpeciesK sp;
    const int32_t *cell_off;
    const double *scratch;  // old/new positions written by launch_particles
    const double *pz, *w, *invgf;
    int64_t cap;
};
void launch_det_deposit(const Geom &g, const DetSpecies *species, int nspecies, double *pj, int *err,
                        hipStream_t st);
// J ghost sum along each dimension (SyncVectorPatch::sumAllComponents) and
// copy of the patch interiors into the global arrays.
void launch_ghost_sum_and_fold(const Geom &g, double *pj, const JPtrs &J, hipStream_t st);
// The x, y, z passes of the ghost sum alone, on the patch-major ghosted J
// array pj ([patch][Jx|Jy|Jz], SyncVectorPatch::sumAllComponents).
void launch_ghost_sum(const Geom &g, double *pj, hipStream_t st);
int64_t patch_j_size(const Geom &g);  // doubles per patch (Jx+Jy+Jz)

// ---- sort (sort.hip) --------------------------------------------------
// Lane-interleaved storage of a cluster.  The n particles of a cluster, in
// cell-major order (rank p = 0..n-1), are dealt into 64 columns of
// consecutive ranks (the first `full` columns hold q = ceil(n/64) ranks, the
// others q-1) and stored row by row: slot = row*64 + column.  A wavefront
// reading 64 consecutive slots then loads coalesced, and its lanes hold
// particles about n/64 ranks apart, i.e. in different cells: their LDS
// stencil accesses spread over the tile instead of piling onto the same
// addresses.  cluster_rank is the inverse map.
__host__ __device__ inline int cluster_slot(int p, int n)
{
    const int q = (n + 63) >> 6, full = n - 64 * (q - 1);
    int col, row;
    if (p < full * q) {
        col = p / q;
        row = p - col * q;
    } else {
        const int r = p - full * q;
        const int k = r / (q - 1);
        col = full + k;
        row = r - k * (q - 1);
    }
    return row * 64 + col;
}
__host__ __device__ inline int cluster_rank(int m, int n)
{
    const int q = (n + 63) >> 6, full = n - 64 * (q - 1);
    const int row = m >> 6, col = m & 63;
    return col < full ? col * q + row : full * q + (col - full) * (q - 1) + row;
}
// Pair-interleaved storage: ranks 2j and 2j+1 share a row in the adjacent
// lanes 2c, 2c+1, the pairs dealt into 32 pair columns the way cluster_slot
// deals ranks into 64 columns (an odd last rank takes the last slot).
// Consecutive ranks mostly lie in the same cell, so the two lanes of a pair
// usually deposit onto the same tile words and can add their contributions
// before one of them issues the atomic; the 32 pairs of a row lie about
// n/32 ranks apart, in different cells.  cluster_rank_pair is the inverse.
__host__ __device__ inline int cluster_slot_pair(int p, int n)
{
    const int m = n >> 1;
    if (p >= 2 * m) return p;
    const int j = p >> 1;
    const int q = (m + 31) >> 5, full = m - 32 * (q - 1);
    int col, row;
    if (j < full * q) {
        col = j / q;
        row = j - col * q;
    } else {
        const int r = j - full * q;
        const int k = r / (q - 1);
        col = full + k;
        row = r - k * (q - 1);
    }
    return row * 64 + 2 * col + (p & 1);
}
__host__ __device__ inline int cluster_rank_pair(int m, int n)
{
    const int np = n >> 1;
    if (m >= 2 * np) return m;
    const int q = (np + 31) >> 5, full = np - 32 * (q - 1);
    const int row = m >> 6, col = (m & 63) >> 1;
    const int j = col < full ? col * q + row : full * q + (col - full) * (q - 1) + row;
    return 2 * j + (m & 1);
}
// Cluster layouts of the sort: 0 cell-major, 1 lane-interleaved
// (cluster_slot), 2 pair-interleaved (cluster_slot_pair).  Slot of the
// cell-major rank p of a cluster of n particles, and the inverse.
__host__ __device__ inline int layout_slot(int lay, int p, int n)
{
    return lay == 1 ? cluster_slot(p, n) : (lay == 2 ? cluster_slot_pair(p, n) : p);
}
__host__ __device__ inline int layout_rank(int lay, int m, int n)
{
    return lay == 1 ? cluster_rank(m, n) : (lay == 2 ? cluster_rank_pair(m, n) : m);
}

struct SortBufs {
    int32_t *counts;   // ncells, kept at zero between calls
    int32_t *cell_off; // ncells + 1
    int32_t *perm;     // cap
    int32_t *tmp;      // cap (cell-major order before the interleaved layout)
    int32_t *rank;     // cap (tie-break of the canonical order of the particles that left their cluster)
    int32_t *big;      // cap / (kFixRegs + 1) + 1, or null: cells too full for the register ranking
    int32_t *nbig;     // 2 (count, queue head of `big`)
    int32_t *bsum;     // scan scratch
    int64_t nbsum;
    // Placed stayers: when the particle kernel is given `stay`, a particle
    // whose destination cell lies in its own cluster gets its rank r among
    // that cluster's stayers of the cell as the key's top five bits when
    // r < kPlaceMax (every other particle: kPlaceAtomic there), and the
    // cluster stores min(stayers, kPlaceMax) of each of its cells here.  The
    // sort puts the placed ones at cell start + r and hands out the rest of
    // the bucket with the atomic count-down; its scan clears this array.
    uint8_t *stay;     // ncells rounded up to a multiple of 4, zero between calls
    int32_t *clus_old; // clusters, or null: each cluster's first slot in the order before the sort
};
constexpr uint32_t kPlaceMax = 31, kPlaceAtomic = 31;
int64_t sort_bsum_size(int64_t ncells);
// Counting sort of particles by destination cell, ties by (direction,
// slot) when in_cell_order (otherwise the order inside a cell is that of
// the atomic scatter): fills cell_off and perm from keys[0..n).  With
// counted = true the histogram of destination cells is already in sb.counts
// (the particle kernel builds it); otherwise it is computed from the keys.
// interleaved: each cluster's range of perm is stored in the lane-
// interleaved layout (cluster_slot) instead of cell-major order; cell_off
// then still gives every cluster's range, the cells' sub-ranges refer to
// cell-major ranks.  The canonical order breaks ties by the source slot, or
// by tie[slot] when given (the slot's cell-major position in the previous
// sorted order, for sources stored interleaved).  placed (with counted and
// not in_cell_order): the keys and sb.stay carry the particle kernel's
// placed stayers (SortBufs::stay).
void launch_sort(const Geom &g, const uint32_t *keys, int64_t n, SortBufs &sb, hipStream_t st,
                 bool counted = false, bool in_cell_order = true, int layout = 0,
                 const int32_t *tie = nullptr, bool placed = false);
// Re-stores the current sorted order in another cluster layout (perm via
// tmp); keys are those the order was sorted from.
void launch_relayout(const Geom &g, const uint32_t *keys, int64_t n, SortBufs &sb, int from_layout,
                     int to_layout, hipStream_t st);
// Keys of freshly loaded particles (positions in A, slot order = input).
void launch_bin_initial(const Geom &g, const PartPtrs &A, int64_t n, uint32_t *keys, hipStream_t st);

// ---- mixed-precision form (particles_mixed.hip) -------------------------
// Same contract as launch_particles (always the atomic tile deposit), with
// particles binned by their nearest grid node instead of their cell: the
// node of particle i is keysA[i] (cell-key layout), its offsets from it are
// in [-1/2, 1/2); its new key goes to keysB[slot].
// One species of a merged particle launch of the mixed form (device array).
struct MixSpecies {
    SpeciesKM k;
    MixPtrs A, B;                   // input / output buffers
    const uint32_t *keysA;          // keys of the input buffer
    uint32_t *keysB;                // keys written for the output buffer
    const int32_t *perm, *cell_off;  // sorted view of the input
    int32_t *counts;                // destination histogram
    int32_t *rank;                  // or null: cell-major input position of the output slots that left their cluster
    uint8_t *stay;                  // or null: placed stayers per cell (SortBufs::stay)
    int lay;                        // cluster layout of the input (layout_slot)
};
// Particle kernel of the mixed form over ns <= kMaxMerged species that share
// one J tile per cluster; cr_max = the largest current contribution of those
// species (sets the fixed-point unit of the tile).  persistent_blocks > 0
// selects the persistent form: that many workgroups pull the non-empty
// clusters from a work list (work: one int per cluster, wcount: 2 ints).
// tile_x: clusters per workgroup along x (the patch-tile extent), 1 or 2.
// paired: the species are stored pair-interleaved (layout 2): order-2
// one-cluster tiles then pre-sum the tile words of same-node lane pairs.
// node_acc: the particles that keep their node are staged per node and
// their tile words summed node by node in registers (one atomic per node
// and word; mixed_node_acc_ok).
void launch_particles_mixed(const Geom &g, const MixSpecies *specs, int ns, double cr_max, const GatherFields &F,
                            const JPtrs &J, int *err, int threads, bool xcd, int tile, int tile_x, bool paired,
                            bool node_acc, int persistent_blocks, int *work, int *wcount, hipStream_t st);
// Whether the mixed kernel has a patch tile of tile_x clusters for this
// geometry and launch form.
bool mixed_tile_x_ok(const Geom &g, int tile_x, bool persistent);
// Whether the node-accumulated deposit runs with this geometry and launch.
bool mixed_node_acc_ok(const Geom &g, int threads, int tile_x, bool paired, bool persistent);
// Resident workgroups of the persistent mixed kernel on the current device.
int mixed_persistent_blocks(const Geom &g, int threads, int tile);
void launch_to_mixed(const Geom &g, const PartPtrs &X, int64_t n, const MixPtrs &M, uint32_t *keys, hipStream_t st);
// FP64 global positions of the mixed particles in buffer order, and the
// key (direction 13) of the cell each lies in.
void launch_from_mixed(const Geom &g, const MixPtrs &M, const uint32_t *keys, int64_t n, const PartPtrs &X,
                       uint32_t *cell_keys, hipStream_t st);
// Particles per patch (counts[npatches]) by the cell each position lies in.
void launch_patch_counts_mixed(const Geom &g, const MixPtrs &M, const uint32_t *keys, int64_t n, int32_t *counts,
                               hipStream_t st);
void launch_ukin_mixed(const MixPtrs &M, const int32_t *perm, int64_t n, double *partials, int nblocks,
                       hipStream_t st);
void launch_rho_mixed(const Geom &g, const SpeciesKM &sp, const MixPtrs &M, const uint32_t *keys, const int32_t *perm,
                      int64_t n, double *rho, hipStream_t st);

// ---- initialisation (init.hip) ---------------------------------------
// Particle i lives in cell cells[i/ppc] (or cell i/ppc when cells == null).
// Writes FP64 particles to A, or (when M.ox != null) mixed particles to M
// plus their cell keys (direction 13) to `keys`.
// Momenta: Gaussian components of std sigma, or, with a table uq[0..nq) of
// |u| quantiles, an isotropic Maxwell-Juttner draw; plus drift along x.
void launch_init_like(const Geom &g, const PartPtrs &S, const MixPtrs &SM, const uint32_t *skeys, const PartPtrs &D,
                      const MixPtrs &DM, uint32_t *dkeys, int64_t n, double wscale, double sigma, const double *uq,
                      int nq, double drift, uint64_t seed, hipStream_t st);
void launch_init(const Geom &g, const PartPtrs &A, const MixPtrs &M, uint32_t *keys, int64_t n, int ppc,
                 const int32_t *cells, const double *weights, double weight, double sigma, const double *uq, int nq,
                 double drift, uint64_t seed, hipStream_t st);

// ---- fields (fields.hip) ----------------------------------------------
struct FieldSet {
    double *E[3], *B[3], *Bm[3], *J[3];
};
// saveMagneticFields + MA + MF + centerMagneticFields on the periodic grid.
void launch_maxwell(const Geom &g, const FieldSet &f, hipStream_t st);
void launch_fill(double *p, int64_t n, double v, hipStream_t st);
void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);
// Diagnostics.
// fids: the SMI_* component ids of the fields (their staggering).
void launch_sumsq(const Geom &g, const double *const *fields, const int *fids, int nf, double *partials,
                  int nblocks, hipStream_t st);
void launch_ukin(const PartPtrs &A, const int32_t *perm, int64_t n, double *partials, int nblocks,
                 hipStream_t st);
// Sum of n weights (FP64 w or FP32 wf), per-block partials.
void launch_wsum(const double *w, const float *wf, int64_t n, double *partials, int nblocks, hipStream_t st);
void launch_rho(const Geom &g, const SpeciesK &sp, const PartPtrs &A, const int32_t *perm, int64_t n,
                double *rho, hipStream_t st);
// Same density added in a fixed order (one thread per node gathering from
// the sorted cells; cell_off / perm / lay as the sort left them).
void launch_rho_gather(const Geom &g, const SpeciesK &sp, const PartPtrs &A, const int32_t *perm,
                       const int32_t *cell_off, int lay, double *rho, hipStream_t st);
void launch_gauss(const Geom &g, const FieldSet &f, double *rho_to_residual, hipStream_t st);

}  // namespace smi
