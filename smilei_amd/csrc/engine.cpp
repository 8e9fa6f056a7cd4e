// engine.cpp -- context API of the MI355X PIC library (host side).
//
// A context owns one simulation on one GPU: the periodic global field
// arrays, the particles of every species in HBM (double-buffered SoA plus
// the sorted view `perm`), the per-patch scratch of the deterministic
// deposit, and one HIP stream.  smi_step() enqueues, per time step, the
// sequence of VectorPatch::dynamics (src/Patch/VectorPatch.cpp:4763),
// sumDensities (:819), solveMaxwell (:965) and finalizeExchParticlesAndSort
// (src/Patch/SyncVectorPatch.cpp:44) as a handful of batched launches, and
// can replay it as a captured hipGraph.
//
// The namelist checks mirror the reference factories: unsupported values
// are rejected with SMI_ERR_INVALID and a message instead of ERROR()+abort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/smilei_amd.h"
#include "kernels.h"

using namespace smi;

namespace {

thread_local std::string g_last_error;

constexpr int kReduceBlocks = 1024;

struct SpeciesState {
    std::string name;
    double mass = 1, charge = -1;
    int64_t cap = 0, n = 0;
    void *buf[2][7] = {};          // x y z ux uy uz w: FP64, or FP32 (offsets) in the mixed form
    uint32_t *keys[2] = {};        // cell keys of the particles of buffer b
    SortBufs sb{};
    int lay = 0;                   // cluster layout of sb.perm (layout_slot: 0 cell-major, 1 lane-, 2 pair-interleaved)
    double *scratch = nullptr;     // 7*cap, deterministic mode
    SpeciesK k{};
    SpeciesKM km{};
    double w_max = 0.;             // largest |weight| (fixed-point range of the mixed form)
    double time_frozen = 0.;       // Species::time_frozen_: immobile while time_dual <= time_frozen
    bool mirrored = false;         // both ping-pong buffers hold the current particles (frozen species)

    PartPtrs view(int b) const
    {
        double *const *q = (double *const *)buf[b];
        return PartPtrs{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
    }
    MixPtrs mview(int b) const
    {
        float *const *q = (float *const *)buf[b];
        return MixPtrs{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
    }
};

template <class T>
T *dalloc(int64_t n)
{
    void *p = nullptr;
    if (n <= 0) n = 1;
    SMI_HIP(hipMalloc(&p, (size_t)n * sizeof(T)));
    return (T *)p;
}

// Temporary device buffer released on scope exit (also on errors).
template <class T>
struct DevBuf {
    T *p = nullptr;
    void alloc(int64_t n) { p = dalloc<T>(n); }
    ~DevBuf() { (void)hipFree(p); }
};

bool streq(const char *a, const char *b) { return a && b && std::strcmp(a, b) == 0; }

}  // namespace

struct smi_context {
    smi_params p{};
    Geom g{};
    hipStream_t st = nullptr;
    double *F[SMI_NFIELDS] = {};
    std::vector<SpeciesState> sp;
    double *pj = nullptr;          // per-patch ghosted J (deterministic mode)
    DetSpecies *d_det = nullptr;   // device copy of the deposit descriptors
    int d_det_n = 0;
    MixSpecies *d_mix = nullptr;   // mixed-form particle-launch descriptors [2][species]
    FpSpecies *d_fp = nullptr;     // FP64-form particle-launch descriptors [2][species]
    double *partials = nullptr;    // reduction scratch
    double *rho = nullptr;         // diagnostics scratch (ncells)
    int *err = nullptr;            // device flag raised by kernels on corrupt particles
    int cur = 0;                   // particle buffer holding the current state
    int64_t steps = 0;             // time steps done (time_dual of the next step = (steps + 1.5) dt)
    std::vector<int> moving;       // species pushed by the next step, in id order (launch descriptors)
    bool mixed = false;            // SMI_MIXED form
    int interleave = 1;            // cluster layout of the sort: 0 cell-major, 1 lane-, 2 pair-interleaved (tuning)
    int threads = 256;             // workgroup size of the particle kernel (tuning)
    int persistent = 0;            // persistent work-list form of the particle kernel (tuning)
    int cell_order = 1;            // canonical order inside a cell (atomic deposit may drop it)
    int xcd = 0;                   // XCD-aware cluster order of the grid launch (tuning)
    int tile = 1;                  // J-tile shape of the mixed kernel: rows padded by one word (tuning)
    int tile_x = 1;                // patch-tile extent of the mixed kernel: clusters along x per workgroup (tuning)
    int node_acc = 0;              // mixed kernel: stayers' tile words summed node by node (tuning)
    int persistent_blocks = 0;     // its grid, from the occupancy of the current knobs
    int *d_work = nullptr;         // work list (one int per cluster) + head/count (2 ints)
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev[6] = {};

    FieldSet fields() const
    {
        FieldSet f;
        for (int c = 0; c < 3; ++c) {
            f.E[c] = F[SMI_EX + c];
            f.B[c] = F[SMI_BX + c];
            f.Bm[c] = F[SMI_BXM + c];
            f.J[c] = F[SMI_JX + c];
        }
        return f;
    }
};

namespace {

void check_ctx(const smi_context *c)
{
    if (!c) throw Error(SMI_ERR_INVALID, "null context");
}

// Cluster layout the tile-deposit kernels (mixed form, atomic FP64 form)
// read: the tuning's (lane-interleaved by default); the deterministic
// deposit walks the cells in order and keeps cell-major order (0).
int want_interleaved(const smi_context *c)
{
    return c->p.deposit != SMI_DEPOSIT_DETERMINISTIC ? c->interleave : 0;
}

// The particle kernel records each input particle's cell-major position
// (the tie-break of the canonical in-cell order) when the input is stored
// interleaved and the next sort keeps the canonical order.
bool want_rank(const smi_context *c)
{
    return want_interleaved(c) != 0 && c->cell_order;
}

// The particle kernel places its cluster's stayers in the step's sort
// (SortBufs::stay) when the order inside cells is the scatter's own.
bool want_placed(const smi_context *c)
{
    return c->p.deposit != SMI_DEPOSIT_DETERMINISTIC && !c->cell_order;
}

// Sort of species S from keys; tie: the kernel's recorded positions (the
// step's sort) or null (the slots themselves: freshly loaded particles).
void sort_species(smi_context *c, SpeciesState &S, const uint32_t *keys, bool counted, bool ordered,
                  const int32_t *tie = nullptr)
{
    if (!counted) S.mirrored = false;  // freshly loaded particles: one buffer holds them
    S.lay = want_interleaved(c);
    launch_sort(c->g, keys, S.n, S.sb, c->st, counted, ordered, S.lay, tie, counted && want_placed(c));
}

void drop_graphs(smi_context *c)
{
    for (auto &gx : c->graph)
        if (gx) {
            (void)hipGraphExecDestroy(gx);
            gx = nullptr;
        }
}

// Device descriptors of the deterministic deposit for both ping-pong states
// (set k: input buffer k, output buffer 1-k), refreshed when species change.
void upload_det(smi_context *c)
{
    if (c->moving.empty()) return;
    std::vector<DetSpecies> v;
    for (int k = 0; k < 2; ++k)
        for (int id : c->moving) {
            const SpeciesState &s = c->sp[id];
            DetSpecies d;
            d.sp = s.k;
            d.cell_off = s.sb.cell_off;
            d.scratch = s.scratch;
            d.pz = (const double *)s.buf[1 - k][5];
            d.w = (const double *)s.buf[1 - k][6];
            d.invgf = nullptr;
            d.cap = s.cap;
            v.push_back(d);
        }
    SMI_HIP(hipMemcpyAsync(c->d_det, v.data(), v.size() * sizeof(DetSpecies), hipMemcpyHostToDevice, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));  // v is released on return
}

// Mixed-form launch descriptors for both ping-pong states (set k: input
// buffer k, output 1-k), refreshed when species change.
void upload_mix(smi_context *c)
{
    if (c->moving.empty() || !c->mixed) return;
    std::vector<MixSpecies> v;
    for (int k = 0; k < 2; ++k)
        for (int id : c->moving) {
            const SpeciesState &s = c->sp[id];
            MixSpecies m;
            m.k = s.km;
            m.A = s.mview(k);
            m.B = s.mview(1 - k);
            m.keysA = s.keys[k];
            m.keysB = s.keys[1 - k];
            m.perm = s.sb.perm;
            m.cell_off = s.sb.cell_off;
            m.counts = s.sb.counts;
            m.rank = want_rank(c) ? s.sb.rank : nullptr;
            m.stay = want_placed(c) ? s.sb.stay : nullptr;
            m.lay = want_interleaved(c);
            v.push_back(m);
        }
    SMI_HIP(hipMemcpyAsync(c->d_mix, v.data(), v.size() * sizeof(MixSpecies), hipMemcpyHostToDevice, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));  // v is released on return
}

// Grid of the persistent particle kernel for the current knobs (queried
// outside any graph capture).
void refresh_persistent(smi_context *c)
{
    if (!c->persistent) c->persistent_blocks = 0;
    else if (c->mixed) c->persistent_blocks = mixed_persistent_blocks(c->g, c->threads, c->tile);
    else c->persistent_blocks = fp_persistent_blocks(c->g, c->p.deposit != SMI_DEPOSIT_DETERMINISTIC, c->threads);
}

// The particle kernel launches of one step reading buffer a, writing 1 - a.
void enqueue_particles(smi_context *c, int a)
{
    const Geom &g = c->g;
    const bool det = c->p.deposit == SMI_DEPOSIT_DETERMINISTIC;
    FieldSet f = c->fields();
    JPtrs J{f.J[0], f.J[1], f.J[2]};
    GatherFields GF{f.E[0], f.E[1], f.E[2], f.Bm[0], f.Bm[1], f.Bm[2]};
    if (c->mixed) {
        // species share the cluster's field and J tiles, kMaxMerged per launch
        const int ns = (int)c->moving.size();
        for (int g0 = 0; g0 < ns; g0 += kMaxMerged) {
            const int n = std::min(kMaxMerged, ns - g0);
            double cr_max = 0.;
            for (int q = g0; q < g0 + n; ++q) cr_max = std::max(cr_max, c->sp[c->moving[q]].km.cr_max);
            // the two-cluster tile has no persistent form: one cluster there
            const int tx = mixed_tile_x_ok(g, c->tile_x, c->persistent != 0) ? c->tile_x : 1;
            const bool paired = want_interleaved(c) == 2;
            // the node-accumulated deposit where the launch allows it, else
            // the per-particle one
            const bool na = c->node_acc && mixed_node_acc_ok(g, c->threads, tx, paired, c->persistent != 0);
            launch_particles_mixed(g, c->d_mix + a * ns + g0, n, cr_max, GF, J, c->err, c->threads, c->xcd != 0, c->tile,
                                   tx, paired, na, c->persistent ? c->persistent_blocks : 0, c->d_work,
                                   c->d_work + g.npatches * g.clpp, c->st);
        }
    } else {
        const int ns = (int)c->moving.size();
        for (int g0 = 0; g0 < ns; g0 += kMaxMerged)
            launch_particles(g, c->d_fp + a * ns + g0, std::min(kMaxMerged, ns - g0), GF, J, !det, c->err,
                             c->threads, c->xcd != 0, c->persistent ? c->persistent_blocks : 0, c->d_work,
                             c->d_work + g.npatches * g.clpp, c->st);
    }
}

// FP64-form launch descriptors for both ping-pong states.
void upload_fp(smi_context *c)
{
    if (c->moving.empty() || c->mixed) return;
    std::vector<FpSpecies> v;
    for (int k = 0; k < 2; ++k)
        for (int id : c->moving) {
            const SpeciesState &s = c->sp[id];
            FpSpecies f;
            f.k = s.k;
            f.A = s.view(k);
            f.B = s.view(1 - k);
            f.perm = s.sb.perm;
            f.cell_off = s.sb.cell_off;
            f.keys = s.keys[1 - k];
            f.counts = s.sb.counts;
            f.scratch = s.scratch;
            f.rank = want_rank(c) ? s.sb.rank : nullptr;
            f.stay = want_placed(c) ? s.sb.stay : nullptr;
            f.lay = want_interleaved(c);
            f.cap = s.cap;
            v.push_back(f);
        }
    SMI_HIP(hipMemcpyAsync(c->d_fp, v.data(), v.size() * sizeof(FpSpecies), hipMemcpyHostToDevice, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));  // v is released on return
}

// After a change of the layout or cell-order knobs: every species' current
// order re-stored in the wanted layout (the order itself is kept), launch
// descriptors refreshed.
void apply_layout(smi_context *c)
{
    const int want = want_interleaved(c);
    for (auto &S : c->sp) {
        launch_relayout(c->g, S.keys[c->cur], S.n, S.sb, S.lay, want, c->st);
        S.lay = want;
    }
    upload_fp(c);
    upload_mix(c);
    SMI_HIP(hipStreamSynchronize(c->st));
}

// One PIC time step reading particle buffer c->cur and writing 1 - c->cur.
void enqueue_step(smi_context *c, bool timed)
{
    const int a = c->cur;
    const Geom &g = c->g;
    const bool det = c->p.deposit == SMI_DEPOSIT_DETERMINISTIC;
    FieldSet f = c->fields();
    JPtrs J{f.J[0], f.J[1], f.J[2]};
    if (timed) SMI_HIP(hipEventRecord(c->ev[0], c->st));
    if (!det)
        for (int q = 0; q < 3; ++q) launch_fill(f.J[q], g.ncells, 0.0, c->st);  // restartRhoJ
    enqueue_particles(c, a);
    if (timed) SMI_HIP(hipEventRecord(c->ev[1], c->st));
    if (det && !c->sp.empty()) {
        // every species' deposit (none: an empty J, still ghost-summed)
        launch_det_deposit(g, c->d_det + a * (int)c->moving.size(), (int)c->moving.size(), c->pj, c->err, c->st);
        launch_ghost_sum_and_fold(g, c->pj, J, c->st);
    }
    if (timed) SMI_HIP(hipEventRecord(c->ev[2], c->st));
    launch_maxwell(g, f, c->st);
    if (timed) SMI_HIP(hipEventRecord(c->ev[3], c->st));
    const bool ordered = det || c->cell_order;
    for (int id : c->moving) {
        SpeciesState &s = c->sp[id];
        sort_species(c, s, s.keys[1 - a], /*counted=*/true, ordered, want_rank(c) ? s.sb.rank : nullptr);
    }
    if (timed) SMI_HIP(hipEventRecord(c->ev[4], c->st));
}

void flip(smi_context *c) { c->cur ^= 1; }

// Species pushed by the step taken after c->steps steps: those with
// time_dual = (steps + 1.5) dt > time_frozen (Species::dynamics,
// Species.cpp:568 and 832-866; time_dual as advanced in Smilei.cpp:487-488).
std::vector<int> moving_now(const smi_context *c)
{
    std::vector<int> m;
    const double time_dual = ((double)c->steps + 1.5) * c->g.dt;
    for (int id = 0; id < (int)c->sp.size(); ++id)
        if (time_dual > c->sp[id].time_frozen) m.push_back(id);
    return m;
}

// Before a step (outside any graph capture): a frozen species is left out
// of the particle launch and the sort, so both ping-pong buffers must hold
// its particles (the buffer index flips every step); the launch descriptors
// follow the set of moving species, and graphs captured for another set are
// dropped.
void prepare_step(smi_context *c, bool force = false)
{
    std::vector<int> mv = moving_now(c);
    for (int id = 0; id < (int)c->sp.size(); ++id) {
        SpeciesState &S = c->sp[id];
        if (S.mirrored || std::find(mv.begin(), mv.end(), id) != mv.end()) continue;
        const size_t esz = c->mixed ? sizeof(float) : sizeof(double);
        const int a = c->cur;
        if (S.n > 0) {
            for (int q = 0; q < 7; ++q)
                SMI_HIP(hipMemcpyAsync(S.buf[1 - a][q], S.buf[a][q], S.n * esz, hipMemcpyDeviceToDevice, c->st));
            SMI_HIP(hipMemcpyAsync(S.keys[1 - a], S.keys[a], S.n * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                   c->st));
        }
        S.mirrored = true;
    }
    if (force || mv != c->moving) {
        c->moving = mv;
        drop_graphs(c);
        upload_det(c);
        upload_fp(c);
        upload_mix(c);
    }
}

// Largest current contribution |q| w_max / V * max(dx/dt) of a species: sets
// the fixed-point unit of the mixed form's J tile.
void set_quantum(const Geom &g, SpeciesState &s, double w_max)
{
    double cell_volume = 1.0, dxdt = 0.;
    for (int d = 0; d < g.nd; ++d) {
        cell_volume *= g.dx[d];
        dxdt = std::max(dxdt, g.dx[d] / g.dt);
    }
    s.w_max = w_max;
    double cmax = std::fabs(s.charge) / cell_volume * w_max * dxdt;
    if (!(cmax > 0)) cmax = 1.0;
    s.km.cr_max = cmax;
}

int fail(const Error &e)
{
    g_last_error = e.what();
    return e.code;
}

}  // namespace

void smi::set_error(const std::string &msg) { g_last_error = msg; }

#define SMI_TRY try {
#define SMI_CATCH                                                    \
    }                                                                \
    catch (const Error &e) { return fail(e); }                       \
    catch (const std::exception &e) { return fail(Error(SMI_ERR_INVALID, e.what())); }

extern "C" {

const char *smi_last_error(void) { return g_last_error.c_str(); }

int smi_create(const smi_params *p, smi_context **out)
{
    SMI_TRY
    if (!p || !out) throw Error(SMI_ERR_INVALID, "null argument");
    *out = nullptr;
    int nd;
    if (streq(p->geometry, "3Dcartesian") || !p->geometry) nd = 3;
    else if (streq(p->geometry, "2Dcartesian")) nd = 2;
    else throw Error(SMI_ERR_INVALID, std::string("unsupported geometry: ") + p->geometry);
    // InterpolatorFactory / ProjectorFactory: order 2 (2D, 3D) or 4 (3D)
    if (!(p->interpolation_order == 2 || (p->interpolation_order == 4 && nd == 3)))
        throw Error(SMI_ERR_INVALID, "interpolation_order must be 2 (2D/3D) or 4 (3D)");
    if (p->interpolator && !streq(p->interpolator, "momentum-conserving"))
        throw Error(SMI_ERR_INVALID, std::string("unsupported interpolator: ") + p->interpolator);
    if (p->pusher && !streq(p->pusher, "boris") && !streq(p->pusher, "Boris"))
        throw Error(SMI_ERR_INVALID, std::string("unsupported pusher: ") + p->pusher);
    if (p->maxwell_solver && !streq(p->maxwell_solver, "Yee"))
        throw Error(SMI_ERR_INVALID, std::string("unsupported maxwell_solver: ") + p->maxwell_solver);
    if (p->precision != SMI_PLAIN_FP64 && p->precision != SMI_MIXED)
        throw Error(SMI_ERR_INVALID, "precision must be SMI_PLAIN_FP64 or SMI_MIXED");
    if (p->precision == SMI_MIXED && nd != 3)
        throw Error(SMI_ERR_INVALID, "the mixed-precision form is available in 3Dcartesian");
    if (p->precision == SMI_MIXED && p->deposit == SMI_DEPOSIT_DETERMINISTIC)
        throw Error(SMI_ERR_INVALID, "the deterministic deposit is a mode of the plain FP64 form");
    if (p->deposit != SMI_DEPOSIT_ATOMIC && p->deposit != SMI_DEPOSIT_DETERMINISTIC)
        throw Error(SMI_ERR_INVALID, "deposit must be SMI_DEPOSIT_ATOMIC or SMI_DEPOSIT_DETERMINISTIC");
    if (!(p->timestep > 0)) throw Error(SMI_ERR_INVALID, "timestep must be > 0");

    Geom g{};
    g.nd = nd;
    g.order = p->interpolation_order;
    g.os = g.order;
    int64_t ncells = 1;
    for (int d = 0; d < 3; ++d) {
        if (d < nd) {
            if (p->n_cells[d] <= 0 || p->patch_size[d] <= 0 || !(p->cell_length[d] > 0))
                throw Error(SMI_ERR_INVALID, "n_cells, patch_size and cell_length must be positive");
            if (p->patch_size[d] % 4) throw Error(SMI_ERR_INVALID, "patch_size must be a multiple of 4");
            if (p->n_cells[d] % p->patch_size[d])
                throw Error(SMI_ERR_INVALID, "patch_size must divide n_cells (Params.cpp:1210)");
            if (p->patch_size[d] < 2 * g.os + 2)
                throw Error(SMI_ERR_INVALID, "patch_size must be >= 2*oversize+2 for the ghost sum");
            g.N[d] = p->n_cells[d];
            g.n[d] = p->patch_size[d];
            g.dx[d] = p->cell_length[d];
        } else {
            g.N[d] = 1;
            g.n[d] = 1;
            g.dx[d] = 1.0;
        }
        g.npatch[d] = g.N[d] / g.n[d];
        g.ncl[d] = d < nd ? g.n[d] / 4 : 1;
        g.dinv[d] = 1.0 / g.dx[d];
        g.plen[d] = (double)g.n[d] * g.dx[d];
        g.xmax[d] = g.dx[d] * (double)g.N[d];
        ncells *= g.N[d];
    }
    for (int d = 0; d < nd; ++d)
        if (!(p->timestep < p->cell_length[d]))
            throw Error(SMI_ERR_INVALID, "timestep must be below every cell_length (c dt < dx): the particle "
                                         "kernels move a particle by at most one cell per step");
    if (ncells >= (1ll << 27)) throw Error(SMI_ERR_INVALID, "at most 2^27 cells per simulation");
    g.ncells = ncells;
    g.cpp = g.n[0] * g.n[1] * g.n[2];
    g.clpp = g.ncl[0] * g.ncl[1] * g.ncl[2];
    g.cells_per_cl = nd == 3 ? 64 : 16;
    g.npatches = g.npatch[0] * g.npatch[1] * g.npatch[2];
    g.dt = p->timestep;
    for (int d = 0; d < 3; ++d) g.dxdt[d] = g.dx[d] / g.dt;

    SMI_HIP(hipSetDevice(p->device));
    auto *c = new smi_context();
    c->p = *p;
    c->g = g;
    c->mixed = p->precision == SMI_MIXED;
    try {
        SMI_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        for (int f = 0; f < SMI_NFIELDS; ++f) {
            c->F[f] = dalloc<double>(ncells);
            launch_fill(c->F[f], ncells, 0.0, c->st);
        }
        c->partials = dalloc<double>(kReduceBlocks);
        c->rho = dalloc<double>(ncells);
        c->err = dalloc<int>(1);
        launch_fill_i32(c->err, 1, 0, c->st);
        if (p->deposit == SMI_DEPOSIT_DETERMINISTIC) c->pj = dalloc<double>(patch_j_size(g) * g.npatches);
        for (auto &e : c->ev) SMI_HIP(hipEventCreate(&e));
        SMI_HIP(hipStreamSynchronize(c->st));
    } catch (...) {
        smi_destroy(c);
        throw;
    }
    *out = c;
    return SMI_OK;
    SMI_CATCH
}

int smi_destroy(smi_context *c)
{
    if (!c) return SMI_OK;
    if (c->st) (void)hipStreamSynchronize(c->st);
    drop_graphs(c);
    for (auto &f : c->F) (void)hipFree(f);
    for (auto &s : c->sp) {
        for (auto &bb : s.buf)
            for (auto &q : bb) (void)hipFree(q);
        for (auto &k : s.keys) (void)hipFree(k);
        (void)hipFree(s.sb.counts);
        (void)hipFree(s.sb.cell_off);
        (void)hipFree(s.sb.perm);
        (void)hipFree(s.sb.tmp);
        (void)hipFree(s.sb.rank);
        (void)hipFree(s.sb.big);
        (void)hipFree(s.sb.nbig);
        (void)hipFree(s.sb.bsum);
        (void)hipFree(s.sb.stay);
        (void)hipFree(s.sb.clus_old);
        (void)hipFree(s.scratch);
    }
    (void)hipFree(c->pj);
    (void)hipFree(c->d_det);
    (void)hipFree(c->d_mix);
    (void)hipFree(c->d_fp);
    (void)hipFree(c->d_work);
    (void)hipFree(c->partials);
    (void)hipFree(c->rho);
    (void)hipFree(c->err);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return SMI_OK;
}

int smi_add_species(smi_context *c, const char *name, double mass, double charge, int64_t capacity, int *id)
{
    SMI_TRY
    check_ctx(c);
    if (!(mass > 0)) throw Error(SMI_ERR_INVALID, "mass must be > 0");
    if (capacity < 0 || capacity >= (1ll << 31)) throw Error(SMI_ERR_INVALID, "capacity must be in [0, 2^31)");
    SMI_HIP(hipSetDevice(c->p.device));
    drop_graphs(c);
    SpeciesState s;
    s.name = name ? name : "";
    s.mass = mass;
    s.charge = charge;
    s.cap = capacity;
    for (auto &bb : s.buf)
        for (auto &q : bb) q = c->mixed ? (void *)dalloc<float>(capacity) : (void *)dalloc<double>(capacity);
    for (auto &k : s.keys) k = dalloc<uint32_t>(capacity);
    s.sb.counts = dalloc<int32_t>(c->g.ncells);
    s.sb.cell_off = dalloc<int32_t>(c->g.ncells + 1);
    s.sb.perm = dalloc<int32_t>(capacity);
    s.sb.tmp = dalloc<int32_t>(capacity);
    s.sb.big = dalloc<int32_t>(capacity / 17 + 1);
    s.sb.nbig = dalloc<int32_t>(2);
    if (c->p.deposit != SMI_DEPOSIT_DETERMINISTIC) s.sb.rank = dalloc<int32_t>(capacity);
    s.sb.nbsum = sort_bsum_size(c->g.ncells);
    s.sb.bsum = dalloc<int32_t>(s.sb.nbsum);
    s.sb.stay = (uint8_t *)dalloc<int32_t>((c->g.ncells + 3) / 4);
    s.sb.clus_old = dalloc<int32_t>(c->g.ncells / c->g.cells_per_cl);
    launch_fill_i32((int32_t *)s.sb.stay, (c->g.ncells + 3) / 4, 0, c->st);
    if (c->p.deposit == SMI_DEPOSIT_DETERMINISTIC) s.scratch = dalloc<double>(7 * capacity);
    launch_fill_i32(s.sb.counts, c->g.ncells, 0, c->st);
    launch_fill_i32(s.sb.cell_off, c->g.ncells + 1, 0, c->st);
    double cell_volume = 1.0;
    for (int d = 0; d < c->g.nd; ++d) cell_volume *= c->g.dx[d];
    const double one_over_mass = 1.0 / mass;
    const double dts2 = c->g.dt / 2.;
    s.k.charge = charge;
    s.k.mass = mass;
    s.k.qmh = charge * one_over_mass * dts2;
    s.k.cw_scale = (1. / cell_volume) * charge;
    s.km.qmh = (float)s.k.qmh;
    s.km.qmh_lo = (float)(s.k.qmh - (double)s.km.qmh);
    s.km.cw_scale = (float)s.k.cw_scale;
    s.km.cwd = s.k.cw_scale;
    for (int d = 0; d < 3; ++d) s.km.dxdt[d] = (float)(c->g.dx[d] / c->g.dt);
    set_quantum(c->g, s, 1.0);
    c->sp.push_back(s);
    c->moving = moving_now(c);
    // deterministic-deposit descriptors for both ping-pong parities
    (void)hipFree(c->d_det);
    c->d_det = nullptr;
    c->d_det_n = (int)c->sp.size();
    c->d_det = dalloc<DetSpecies>(2 * c->d_det_n);
    upload_det(c);
    (void)hipFree(c->d_mix);
    c->d_mix = nullptr;
    (void)hipFree(c->d_fp);
    c->d_fp = nullptr;
    if (!c->mixed) {
        c->d_fp = dalloc<FpSpecies>(2 * c->sp.size());
        upload_fp(c);
    }
    if (c->mixed) {
        c->d_mix = dalloc<MixSpecies>(2 * c->sp.size());
        upload_mix(c);
    }
    if (!c->d_work) c->d_work = dalloc<int>((int64_t)c->g.npatches * c->g.clpp + 2);
    SMI_HIP(hipStreamSynchronize(c->st));
    if (id) *id = (int)c->sp.size() - 1;
    return SMI_OK;
    SMI_CATCH
}

int smi_set_particles(smi_context *c, int s, int64_t n, const double *x, const double *y, const double *z,
                      const double *px, const double *py, const double *pz, const double *w)
{
    SMI_TRY
    check_ctx(c);
    if (s < 0 || s >= (int)c->sp.size()) throw Error(SMI_ERR_INVALID, "bad species id");
    SpeciesState &S = c->sp[s];
    if (n < 0 || n > S.cap) throw Error(SMI_ERR_CAPACITY, "particle count exceeds the species capacity");
    const double *src[7] = {x, y, z, px, py, pz, w};
    // z may be NULL in 2D; every other array is required when n > 0
    for (int q = 0; q < 7; ++q)
        if (n > 0 && !src[q] && !(q == 2 && c->g.nd < 3)) throw Error(SMI_ERR_INVALID, "null particle array");
    for (int d = 0; d < c->g.nd; ++d) {
        const double *v = src[d];
        for (int64_t i = 0; i < n; ++i)
            if (!(v[i] >= 0. && v[i] < c->g.xmax[d]))
                throw Error(SMI_ERR_INVALID, "particle position outside the simulation box");
    }
    // a non-finite momentum would move its particle by a non-finite offset
    // and spread NaN through J; a non-finite weight also sets the mixed
    // form's fixed-point unit
    for (int q = 3; q < 7; ++q)
        for (int64_t i = 0; i < n; ++i)
            if (!std::isfinite(src[q][i])) throw Error(SMI_ERR_INVALID, "non-finite particle momentum or weight");
    SMI_HIP(hipSetDevice(c->p.device));
    drop_graphs(c);
    SMI_HIP(hipStreamSynchronize(c->st));
    S.n = n;
    if (c->mixed) {
        double wmax = 0.;
        for (int64_t i = 0; i < n; ++i) wmax = std::max(wmax, std::fabs(w[i]));
        set_quantum(c->g, S, wmax);
        // stage FP64 on the device, convert to cell + FP32 offset (the
        // staging buffers are freed on every exit path)
        DevBuf<double> tmp[7];
        for (int q = 0; q < 7; ++q) {
            tmp[q].alloc(n);
            if (n) SMI_HIP(hipMemcpyAsync(tmp[q].p, src[q], n * sizeof(double), hipMemcpyHostToDevice, c->st));
        }
        launch_to_mixed(c->g, PartPtrs{tmp[0].p, tmp[1].p, tmp[2].p, tmp[3].p, tmp[4].p, tmp[5].p, tmp[6].p}, n,
                        S.mview(c->cur), S.keys[c->cur], c->st);
        SMI_HIP(hipStreamSynchronize(c->st));
    } else {
        for (int q = 0; q < 7; ++q) {
            double *dst = (double *)S.buf[c->cur][q];
            if (q == 2 && c->g.nd < 3) {
                launch_fill(dst, n, 0.0, c->st);
                continue;
            }
            if (n) SMI_HIP(hipMemcpyAsync(dst, src[q], n * sizeof(double), hipMemcpyHostToDevice, c->st));
        }
        launch_bin_initial(c->g, S.view(c->cur), n, S.keys[c->cur], c->st);
    }
    sort_species(c, S, S.keys[c->cur], false, true);
    SMI_HIP(hipStreamSynchronize(c->st));
    return SMI_OK;
    SMI_CATCH
}

extern "C++" {  // host helpers of the entry points below, C++ linkage
namespace {

// Quantiles of |u| = |gamma beta| of the isotropic Maxwell-Juttner
// distribution f(u) ~ u^2 exp(-(gamma - 1) / theta), theta = T / (m c^2):
// the CDF integrated (Simpson) on a fine grid up to gamma - 1 = 60 theta
// (tail below e^-50), inverted at nq equally spaced probabilities.
std::vector<double> mj_quantiles(double theta, int nq)
{
    const int nf = 1 << 16;  // even number of intervals
    const double gm = 1.0 + 60.0 * theta, umax = std::sqrt(gm * gm - 1.0);
    std::vector<double> u(nf + 1), f(nf + 1), cdf(nf + 1, 0.0);
    for (int i = 0; i <= nf; ++i) {
        u[i] = umax * i / nf;
        const double gam1 = u[i] * u[i] / (std::sqrt(1.0 + u[i] * u[i]) + 1.0);  // gamma - 1 without cancellation
        f[i] = u[i] * u[i] * std::exp(-gam1 / theta);
    }
    const double h = umax / nf;
    for (int i = 2; i <= nf; i += 2) {
        cdf[i] = cdf[i - 2] + h / 3.0 * (f[i - 2] + 4.0 * f[i - 1] + f[i]);
        cdf[i - 1] = cdf[i - 2] + h / 2.0 * (f[i - 2] + f[i - 1]);  // trapezoid at the odd point
    }
    const double tot = cdf[nf];
    std::vector<double> q(nq);
    int j = 0;
    for (int k = 0; k < nq; ++k) {
        const double target = tot * k / (nq - 1);
        while (j < nf && cdf[j + 1] < target) ++j;
        if (j >= nf) {
            q[k] = umax;
            continue;
        }
        const double span = cdf[j + 1] - cdf[j];
        q[k] = u[j] + (span > 0 ? (target - cdf[j]) / span : 0.0) * h;
    }
    q[0] = 0.0;
    q[nq - 1] = umax;
    return q;
}

constexpr int kMjQuantiles = 4096;

// Momentum distribution of an initialiser: Gaussian components of std sigma,
// or Maxwell-Juttner of temperature theta (m c^2 units) when theta > 0.
struct Momenta {
    double sigma = 0., theta = 0., drift = 0.;
};

// Device table of the Maxwell-Juttner quantiles (or nothing for Gaussian).
struct MjTable {
    DevBuf<double> buf;
    int n = 0;
    explicit MjTable(const Momenta &m)
    {
        if (!(m.theta > 0)) return;
        const std::vector<double> q = mj_quantiles(m.theta, kMjQuantiles);
        buf.alloc(kMjQuantiles);
        SMI_HIP(hipMemcpy(buf.p, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
        n = kMjQuantiles;
    }
};

// Shared tail of the initialisers: draw n particles (particle i in cell
// cells[i / ppc]), bin and sort them.  cells/weights are device arrays or
// null; w_max bounds |weight| (fixed-point range of the mixed form).
void init_particles(smi_context *c, int s, int64_t n, int ppc, const int32_t *d_cells, const double *d_weights,
                    double weight, double w_max, const Momenta &mom, uint64_t seed)
{
    SpeciesState &S = c->sp[s];
    const Geom &g = c->g;
    const MixPtrs none{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (c->mixed) set_quantum(g, S, w_max);
    MjTable mj(mom);
    launch_init(g, S.view(c->cur), c->mixed ? S.mview(c->cur) : none, S.keys[c->cur], n, ppc, d_cells, d_weights,
                weight, mom.sigma, mj.n ? mj.buf.p : nullptr, mj.n, mom.drift, seed * 0x100000001B3ull + (uint64_t)s,
                c->st);
    S.n = n;
    if (!c->mixed) launch_bin_initial(g, S.view(c->cur), n, S.keys[c->cur], c->st);
    sort_species(c, S, S.keys[c->cur], false, true);
    SMI_HIP(hipStreamSynchronize(c->st));
}

double cell_volume_of(const Geom &g)
{
    double v = 1.0;
    for (int d = 0; d < g.nd; ++d) v *= g.dx[d];
    return v;
}

}  // namespace
}  // extern "C++"

namespace {

void init_constant(smi_context *c, int s, int ppc, double density, const Momenta &mom, double radius, uint64_t seed)
{
    if (s < 0 || s >= (int)c->sp.size()) throw Error(SMI_ERR_INVALID, "bad species id");
    if (ppc <= 0) throw Error(SMI_ERR_INVALID, "ppc must be > 0");
    SpeciesState &S = c->sp[s];
    const Geom &g = c->g;
    SMI_HIP(hipSetDevice(c->p.device));
    drop_graphs(c);
    SMI_HIP(hipStreamSynchronize(c->st));
    std::vector<int32_t> act;
    int64_t ncell_act = g.ncells;
    if (radius > 0) {
        // cells whose centre lies within `radius` cells of the transverse axis
        const double cy = 0.5 * g.N[1], cz = 0.5 * g.N[2];
        for (int64_t t = 0; t < g.ncells; ++t) {
            const int k = (int)(t % g.N[2]), j = (int)((t / g.N[2]) % g.N[1]);
            const double ry = j + 0.5 - cy, rz = g.nd == 3 ? k + 0.5 - cz : 0.;
            if (ry * ry + rz * rz <= radius * radius) act.push_back((int32_t)t);
        }
        ncell_act = (int64_t)act.size();
    }
    const int64_t n = ncell_act * ppc;
    if (n > S.cap) throw Error(SMI_ERR_CAPACITY, "initial particles exceed the species capacity");
    DevBuf<int32_t> d_cells;
    if (radius > 0 && ncell_act) {
        d_cells.alloc(ncell_act);
        SMI_HIP(hipMemcpy(d_cells.p, act.data(), act.size() * 4, hipMemcpyHostToDevice));
    }
    const double weight = density * cell_volume_of(g) / ppc;
    init_particles(c, s, n, ppc, radius > 0 ? d_cells.p : nullptr, nullptr, weight, std::fabs(weight), mom, seed);
}

}  // namespace

int smi_init_species(smi_context *c, int s, int ppc, double density, double sigma, double drift_ux, double radius,
                     uint64_t seed)
{
    SMI_TRY
    check_ctx(c);
    Momenta m;
    m.sigma = sigma;
    m.drift = drift_ux;
    init_constant(c, s, ppc, density, m, radius, seed);
    return SMI_OK;
    SMI_CATCH
}

int smi_init_species_mj(smi_context *c, int s, int ppc, double density, double theta, double drift_ux, double radius,
                        uint64_t seed)
{
    SMI_TRY
    check_ctx(c);
    if (!(theta >= 0) || !std::isfinite(theta)) throw Error(SMI_ERR_INVALID, "theta must be finite and >= 0");
    Momenta m;
    m.theta = theta;
    m.drift = drift_ux;
    init_constant(c, s, ppc, density, m, radius, seed);
    return SMI_OK;
    SMI_CATCH
}

int smi_init_species_like(smi_context *c, int s, int src, double weight_scale, double theta, double drift_ux,
                          uint64_t seed)
{
    SMI_TRY
    check_ctx(c);
    if (s < 0 || s >= (int)c->sp.size() || src < 0 || src >= (int)c->sp.size())
        throw Error(SMI_ERR_INVALID, "bad species id");
    if (s == src) throw Error(SMI_ERR_INVALID, "a species cannot take its positions from itself");
    if (!std::isfinite(weight_scale)) throw Error(SMI_ERR_INVALID, "weight_scale must be finite");
    if (!(theta >= 0) || !std::isfinite(theta)) throw Error(SMI_ERR_INVALID, "theta must be finite and >= 0");
    SpeciesState &S = c->sp[s];
    const SpeciesState &R = c->sp[src];
    if (R.n > S.cap) throw Error(SMI_ERR_CAPACITY, "the source species has more particles than this capacity");
    SMI_HIP(hipSetDevice(c->p.device));
    drop_graphs(c);
    SMI_HIP(hipStreamSynchronize(c->st));
    const MixPtrs none{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const PartPtrs pnone{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (c->mixed) set_quantum(c->g, S, R.w_max * std::fabs(weight_scale));
    Momenta mom;
    mom.theta = theta;
    MjTable mj(mom);
    launch_init_like(c->g, c->mixed ? pnone : R.view(c->cur), c->mixed ? R.mview(c->cur) : none, R.keys[c->cur],
                     c->mixed ? pnone : S.view(c->cur), c->mixed ? S.mview(c->cur) : none, S.keys[c->cur], R.n,
                     weight_scale, 0.0, mj.n ? mj.buf.p : nullptr, mj.n, drift_ux,
                     seed * 0x100000001B3ull + (uint64_t)s, c->st);
    S.n = R.n;
    sort_species(c, S, S.keys[c->cur], false, true);
    SMI_HIP(hipStreamSynchronize(c->st));
    return SMI_OK;
    SMI_CATCH
}

namespace {

void init_profile(smi_context *c, int s, const double *density, const double *ppc, const Momenta &mom, uint64_t seed)
{
    if (s < 0 || s >= (int)c->sp.size()) throw Error(SMI_ERR_INVALID, "bad species id");
    if (!density || !ppc) throw Error(SMI_ERR_INVALID, "density and ppc arrays are required");
    SpeciesState &S = c->sp[s];
    const Geom &g = c->g;
    const double vol = cell_volume_of(g);
    const double inv_nd = 1.0 / g.nd;
    // number of particles per cell: ParticleCreator.cpp:217-246 (floor, plus
    // one on a regular lattice of cells for the fractional part; none where
    // the density vanishes); weight = density * V / nPart (createWeight :936)
    std::vector<int32_t> cells;
    std::vector<double> wts;
    double w_max = 0.;
    for (int64_t t = 0; t < g.ncells; ++t) {
        const int64_t idx[3] = {t / ((int64_t)g.N[1] * g.N[2]), (t / g.N[2]) % g.N[1], t % g.N[2]};
        const double nppc = ppc[t];
        if (!(nppc >= 0.) || !std::isfinite(nppc)) throw Error(SMI_ERR_INVALID, "ppc must be finite and >= 0");
        if (!std::isfinite(density[t])) throw Error(SMI_ERR_INVALID, "density must be finite");
        double np = std::floor(nppc);
        double ip;
        if (std::modf(nppc, &ip) > 0) {
            const double rem = std::pow(nppc - std::floor(nppc), -inv_nd);
            bool extra = true;
            for (int d = 0; d < g.nd; ++d) extra = extra && std::fmod((double)idx[d], rem) < 1.;
            if (extra) np += 1;
        }
        double dens = std::fabs(density[t]) < 1e-200 ? 0. : density[t];
        if (np <= 0 || dens == 0.) continue;
        const int npart = (int)np;
        const double w = std::fabs(dens) * vol / npart;
        w_max = std::max(w_max, w);
        for (int q = 0; q < npart; ++q) {
            cells.push_back((int32_t)t);
            wts.push_back(w);
        }
    }
    const int64_t n = (int64_t)cells.size();
    if (n > S.cap) throw Error(SMI_ERR_CAPACITY, "initial particles exceed the species capacity");
    SMI_HIP(hipSetDevice(c->p.device));
    drop_graphs(c);
    SMI_HIP(hipStreamSynchronize(c->st));
    DevBuf<int32_t> d_cells;
    DevBuf<double> d_w;
    if (n) {
        d_cells.alloc(n);
        d_w.alloc(n);
        SMI_HIP(hipMemcpy(d_cells.p, cells.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
        SMI_HIP(hipMemcpy(d_w.p, wts.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    init_particles(c, s, n, 1, d_cells.p, d_w.p, 0., w_max, mom, seed);
}

}  // namespace

int smi_init_species_profile(smi_context *c, int s, const double *density, const double *ppc, double sigma,
                             double drift_ux, uint64_t seed)
{
    SMI_TRY
    check_ctx(c);
    Momenta m;
    m.sigma = sigma;
    m.drift = drift_ux;
    init_profile(c, s, density, ppc, m, seed);
    return SMI_OK;
    SMI_CATCH
}

int smi_init_species_profile_mj(smi_context *c, int s, const double *density, const double *ppc, double theta,
                                double drift_ux, uint64_t seed)
{
    SMI_TRY
    check_ctx(c);
    if (!(theta >= 0) || !std::isfinite(theta)) throw Error(SMI_ERR_INVALID, "theta must be finite and >= 0");
    Momenta m;
    m.theta = theta;
    m.drift = drift_ux;
    init_profile(c, s, density, ppc, m, seed);
    return SMI_OK;
    SMI_CATCH
}

int64_t smi_num_particles(const smi_context *c, int s)
{
    if (!c || s < 0 || s >= (int)c->sp.size()) return -1;
    return c->sp[s].n;
}

int smi_get_particles(smi_context *c, int s, double *x, double *y, double *z, double *px, double *py, double *pz,
                      double *w, int32_t *patch, int32_t *key)
{
    SMI_TRY
    check_ctx(c);
    if (s < 0 || s >= (int)c->sp.size()) throw Error(SMI_ERR_INVALID, "bad species id");
    SpeciesState &S = c->sp[s];
    const int64_t n = S.n;
    SMI_HIP(hipSetDevice(c->p.device));
    SMI_HIP(hipStreamSynchronize(c->st));
    std::vector<int32_t> perm(n);
    std::vector<uint32_t> keys(n);
    std::vector<double> tmp(n);
    double *dst[7] = {x, y, z, px, py, pz, w};
    if (c->mixed) {
        // FP64 positions and the keys of the cells they lie in (the mixed
        // form bins by node), then a counting sort by cell: cell-major order
        DevBuf<double> t[7];
        DevBuf<uint32_t> ck;
        for (auto &q : t) q.alloc(n);
        ck.alloc(n);
        launch_from_mixed(c->g, S.mview(c->cur), S.keys[c->cur], n,
                          PartPtrs{t[0].p, t[1].p, t[2].p, t[3].p, t[4].p, t[5].p, t[6].p}, ck.p, c->st);
        DevBuf<int32_t> cnt, off, pm, bsum;
        cnt.alloc(c->g.ncells);
        off.alloc(c->g.ncells + 1);
        pm.alloc(n);
        SortBufs tb{};
        tb.counts = cnt.p;
        tb.cell_off = off.p;
        tb.perm = pm.p;
        tb.nbsum = sort_bsum_size(c->g.ncells);
        bsum.alloc(tb.nbsum);
        tb.bsum = bsum.p;
        launch_fill_i32(cnt.p, c->g.ncells, 0, c->st);
        launch_sort(c->g, ck.p, n, tb, c->st, false, false, false, nullptr);  // order inside cells: below
        SMI_HIP(hipStreamSynchronize(c->st));
        if (n) {
            SMI_HIP(hipMemcpy(perm.data(), pm.p, n * 4, hipMemcpyDeviceToHost));
            SMI_HIP(hipMemcpy(keys.data(), ck.p, n * 4, hipMemcpyDeviceToHost));
            // inside a cell, by position (x, then y, then z): an order that
            // depends on the state only, not on the buffer history
            std::vector<double> px(n), py(n), pz(n);
            SMI_HIP(hipMemcpy(px.data(), t[0].p, n * sizeof(double), hipMemcpyDeviceToHost));
            SMI_HIP(hipMemcpy(py.data(), t[1].p, n * sizeof(double), hipMemcpyDeviceToHost));
            SMI_HIP(hipMemcpy(pz.data(), t[2].p, n * sizeof(double), hipMemcpyDeviceToHost));
            const uint32_t m27 = (1u << 27) - 1;
            for (int64_t a = 0; a < n;) {
                int64_t b = a + 1;
                while (b < n && (keys[perm[b]] & m27) == (keys[perm[a]] & m27)) ++b;
                if (b - a > 1)
                    std::sort(perm.begin() + a, perm.begin() + b, [&](int32_t u, int32_t v) {
                        if (px[u] != px[v]) return px[u] < px[v];
                        if (py[u] != py[v]) return py[u] < py[v];
                        return pz[u] < pz[v];
                    });
                a = b;
            }
        }
        for (int q = 0; q < 7; ++q) {
            if (!dst[q] || !n) continue;
            SMI_HIP(hipMemcpy(tmp.data(), t[q].p, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < n; ++j) dst[q][j] = tmp[perm[j]];
        }
    } else if (n) {
        SMI_HIP(hipMemcpy(perm.data(), S.sb.perm, n * 4, hipMemcpyDeviceToHost));
        SMI_HIP(hipMemcpy(keys.data(), S.keys[c->cur], n * 4, hipMemcpyDeviceToHost));
    }
    if (!c->mixed && S.lay && n) {
        // back to cell-major order: rank p of a cluster sits at slot layout_slot(p)
        std::vector<int32_t> off(c->g.ncells + 1), lay(perm);
        SMI_HIP(hipMemcpy(off.data(), S.sb.cell_off, off.size() * 4, hipMemcpyDeviceToHost));
        const int cpc = c->g.cells_per_cl;
        for (int64_t cb = 0; cb < c->g.ncells; cb += cpc) {
            const int a = off[cb], m = off[cb + cpc] - a;
            for (int p = 0; p < m; ++p) perm[a + p] = lay[a + layout_slot(S.lay, p, m)];
        }
    }
    if (!c->mixed) {
        for (int q = 0; q < 7; ++q) {
            if (!dst[q]) continue;
            if (q == 2 && c->g.nd < 3) {
                std::fill(dst[q], dst[q] + n, 0.0);
                continue;
            }
            if (n) SMI_HIP(hipMemcpy(tmp.data(), S.buf[c->cur][q], n * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < n; ++j) dst[q][j] = tmp[perm[j]];
        }
    }
    const uint32_t mask = (1u << 27) - 1;
    for (int64_t j = 0; j < n; ++j) {
        const uint32_t gk = keys[perm[j]] & mask;
        if (patch) patch[j] = (int32_t)(gk / (uint32_t)c->g.cpp);
        if (key) key[j] = (int32_t)(gk % (uint32_t)c->g.cpp);
    }
    return SMI_OK;
    SMI_CATCH
}

int smi_patch_counts(smi_context *c, int s, int32_t *counts)
{
    SMI_TRY
    check_ctx(c);
    if (s < 0 || s >= (int)c->sp.size()) throw Error(SMI_ERR_INVALID, "bad species id");
    SMI_HIP(hipSetDevice(c->p.device));
    SMI_HIP(hipStreamSynchronize(c->st));
    if (c->mixed) {  // binned by node: count by the cells the positions lie in
        DevBuf<int32_t> d;
        d.alloc(c->g.npatches);
        const SpeciesState &S = c->sp[s];
        launch_patch_counts_mixed(c->g, S.mview(c->cur), S.keys[c->cur], S.n, d.p, c->st);
        SMI_HIP(hipStreamSynchronize(c->st));
        SMI_HIP(hipMemcpy(counts, d.p, c->g.npatches * sizeof(int32_t), hipMemcpyDeviceToHost));
        return SMI_OK;
    }
    std::vector<int32_t> off(c->g.ncells + 1);
    SMI_HIP(hipMemcpy(off.data(), c->sp[s].sb.cell_off, off.size() * 4, hipMemcpyDeviceToHost));
    for (int p = 0; p < c->g.npatches; ++p)
        counts[p] = off[(int64_t)(p + 1) * c->g.cpp] - off[(int64_t)p * c->g.cpp];
    return SMI_OK;
    SMI_CATCH
}

int smi_set_field(smi_context *c, int f, const double *gl)
{
    SMI_TRY
    check_ctx(c);
    if (f < 0 || f >= SMI_NFIELDS || !gl) throw Error(SMI_ERR_INVALID, "bad field id");
    SMI_HIP(hipSetDevice(c->p.device));
    SMI_HIP(hipStreamSynchronize(c->st));
    SMI_HIP(hipMemcpy(c->F[f], gl, c->g.ncells * sizeof(double), hipMemcpyHostToDevice));
    return SMI_OK;
    SMI_CATCH
}

int smi_get_field(smi_context *c, int f, double *gl)
{
    SMI_TRY
    check_ctx(c);
    if (f < 0 || f >= SMI_NFIELDS || !gl) throw Error(SMI_ERR_INVALID, "bad field id");
    SMI_HIP(hipSetDevice(c->p.device));
    SMI_HIP(hipStreamSynchronize(c->st));
    SMI_HIP(hipMemcpy(gl, c->F[f], c->g.ncells * sizeof(double), hipMemcpyDeviceToHost));
    return SMI_OK;
    SMI_CATCH
}

int smi_step(smi_context *c, int nsteps)
{
    SMI_TRY
    check_ctx(c);
    if (nsteps < 0) throw Error(SMI_ERR_INVALID, "nsteps < 0");
    SMI_HIP(hipSetDevice(c->p.device));
    for (int it = 0; it < nsteps; ++it) {
        prepare_step(c);
        if (c->p.use_graph) {
            // graph[k]: one step reading particle buffer k (captured once)
            const int k = c->cur;
            if (!c->graph[k]) {
                hipGraph_t gr = nullptr;
                SMI_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_step(c, false);
                } catch (...) {
                    // leave the stream out of capture mode, drop the partial graph
                    hipGraph_t partial = nullptr;
                    (void)hipStreamEndCapture(c->st, &partial);
                    if (partial) (void)hipGraphDestroy(partial);
                    throw;
                }
                SMI_HIP(hipStreamEndCapture(c->st, &gr));
                SMI_HIP(hipGraphInstantiate(&c->graph[k], gr, nullptr, nullptr, 0));
                SMI_HIP(hipGraphDestroy(gr));
            }
            SMI_HIP(hipGraphLaunch(c->graph[k], c->st));
        } else {
            enqueue_step(c, false);
        }
        flip(c);
        ++c->steps;
    }
    return SMI_OK;
    SMI_CATCH
}

int smi_set_time_frozen(smi_context *c, int s, double time_frozen)
{
    SMI_TRY
    check_ctx(c);
    if (s < 0 || s >= (int)c->sp.size()) throw Error(SMI_ERR_INVALID, "bad species id");
    if (!(time_frozen == time_frozen)) throw Error(SMI_ERR_INVALID, "time_frozen is NaN");
    c->sp[s].time_frozen = time_frozen;
    return SMI_OK;
    SMI_CATCH
}

int smi_set_step_count(smi_context *c, int64_t steps_done)
{
    SMI_TRY
    check_ctx(c);
    if (steps_done < 0) throw Error(SMI_ERR_INVALID, "steps_done < 0");
    c->steps = steps_done;
    return SMI_OK;
    SMI_CATCH
}

int64_t smi_step_count(const smi_context *c) { return c ? c->steps : -1; }

int smi_synchronize(smi_context *c)
{
    SMI_TRY
    check_ctx(c);
    SMI_HIP(hipStreamSynchronize(c->st));
    int e = 0;
    SMI_HIP(hipMemcpy(&e, c->err, sizeof(int), hipMemcpyDeviceToHost));
    if (e)
        throw Error(SMI_ERR_STATE, "a particle left its cluster stencil (moved more than one cell per step, "
                                   "non-finite position or corrupt input); the step skipped it");
    return SMI_OK;
    SMI_CATCH
}

int smi_profile_step(smi_context *c, float *ms)
{
    SMI_TRY
    check_ctx(c);
    SMI_HIP(hipSetDevice(c->p.device));
    prepare_step(c);
    enqueue_step(c, true);
    flip(c);
    ++c->steps;
    SMI_HIP(hipEventSynchronize(c->ev[4]));
    for (int k = 0; k < 4; ++k) SMI_HIP(hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]));
    // ms[0] particles, ms[1] deterministic J reduction, ms[2] Maxwell, ms[3] sort
    const float part = ms[0], jred = ms[1], maxw = ms[2], sort = ms[3];
    ms[0] = part;
    ms[1] = sort;
    ms[2] = jred;
    ms[3] = maxw;
    return SMI_OK;
    SMI_CATCH
}

void *smi_stream(smi_context *c) { return c ? (void *)c->st : nullptr; }

int smi_set_tuning(smi_context *c, int key, int value)
{
    SMI_TRY
    check_ctx(c);
    if (key == SMI_TUNE_INTERLEAVE) {
        if (value < 0 || value > 2) throw Error(SMI_ERR_INVALID, "SMI_TUNE_INTERLEAVE must be 0, 1 or 2");
        if (value == 2 && !c->mixed)
            throw Error(SMI_ERR_INVALID, "SMI_TUNE_INTERLEAVE = 2 (pair-interleaved) is a mixed-form layout");
        c->interleave = value;
    } else if (key == SMI_TUNE_THREADS) {
        if (value != 64 && value != 128 && value != 256 && value != 512)
            throw Error(SMI_ERR_INVALID, "SMI_TUNE_THREADS must be 64, 128, 256 or 512");
        if (value > particle_max_threads(c->g, c->mixed))
            throw Error(SMI_ERR_INVALID, "SMI_TUNE_THREADS above the particle kernel's limit (256 for the 3D "
                                         "order-4 FP64 kernel)");
        c->threads = value;
    } else if (key == SMI_TUNE_CELL_ORDER) {
        c->cell_order = value != 0;
    } else if (key == SMI_TUNE_XCD) {
        c->xcd = value != 0;
    } else if (key == SMI_TUNE_PERSISTENT) {
        c->persistent = value != 0;
    } else if (key == SMI_TUNE_TILE) {
        if (value != 0 && value != 1) throw Error(SMI_ERR_INVALID, "SMI_TUNE_TILE must be 0 or 1");
        c->tile = value;
    } else if (key == SMI_TUNE_TILE_X) {
        if (value != 1 && value != 2) throw Error(SMI_ERR_INVALID, "SMI_TUNE_TILE_X must be 1 or 2");
        if (value == 2 && !(c->mixed && mixed_tile_x_ok(c->g, 2, false)))
            throw Error(SMI_ERR_INVALID, "SMI_TUNE_TILE_X = 2 needs the mixed form, order 2 and patches of a "
                                         "multiple of 8 cells along x");
        c->tile_x = value;
    } else if (key == SMI_TUNE_NODE_ACC) {
        if (value != 0 && value != 1) throw Error(SMI_ERR_INVALID, "SMI_TUNE_NODE_ACC must be 0 or 1");
        if (value == 1 && !(c->mixed && c->g.order == 2))
            throw Error(SMI_ERR_INVALID, "SMI_TUNE_NODE_ACC = 1 needs the mixed form at order 2");
        c->node_acc = value;
    } else {
        throw Error(SMI_ERR_INVALID, "unknown tuning key");
    }
    drop_graphs(c);
    refresh_persistent(c);
    SMI_HIP(hipSetDevice(c->p.device));
    apply_layout(c);
    return SMI_OK;
    SMI_CATCH
}

int smi_get_tuning(smi_context *c, int key, int *value)
{
    SMI_TRY
    check_ctx(c);
    if (!value) throw Error(SMI_ERR_INVALID, "null output");
    if (key == SMI_TUNE_INTERLEAVE) *value = c->interleave;
    else if (key == SMI_TUNE_THREADS) *value = c->threads;
    else if (key == SMI_TUNE_PERSISTENT) *value = c->persistent;
    else if (key == SMI_TUNE_XCD) *value = c->xcd;
    else if (key == SMI_TUNE_TILE) *value = c->tile;
    else if (key == SMI_TUNE_TILE_X) *value = c->tile_x;
    else if (key == SMI_TUNE_NODE_ACC) *value = c->node_acc;
    else if (key == SMI_TUNE_CELL_ORDER) *value = c->cell_order || c->p.deposit == SMI_DEPOSIT_DETERMINISTIC;
    else throw Error(SMI_ERR_INVALID, "unknown tuning key");
    return SMI_OK;
    SMI_CATCH
}

int smi_autotune(smi_context *c, int reps, float *best_ms)
{
    SMI_TRY
    check_ctx(c);
    if (reps < 1) throw Error(SMI_ERR_INVALID, "reps must be >= 1");
    SMI_HIP(hipSetDevice(c->p.device));
    prepare_step(c);
    SMI_HIP(hipStreamSynchronize(c->st));
    drop_graphs(c);
    // candidates: cluster layout (tile-deposit forms) x J-tile shape (mixed)
    // x deposit by particle or by node (mixed, order 2)
    // x patch-tile extent (mixed, grid launch: one or two clusters along x)
    // x persistent form x cluster order x workgroup size; each timed on the live
    // particle state with the particle launches of one step.  The trial
    // launches only write the output ping-pong buffer (rewritten by the next
    // step), the destination histograms and placed-stayer counts (cleared
    // below) and, in the atomic
    // deposit, J (cleared at the start of every step).  A layout is tried
    // after re-sorting the current state into it.
    const int maps[3] = {1, 0, 2};
    const bool det = c->p.deposit == SMI_DEPOSIT_DETERMINISTIC;
    // the pair-interleaved layout pays in the mixed kernel only (its
    // pre-summed pair deposits)
    const int nmap = det ? 1 : (c->mixed ? 3 : 2), npers = 2, ntile = c->mixed ? 2 : 1;
    const int sizes[3] = {128, 256, 512};
    int best_map = c->interleave, best_threads = c->threads, best_pers = c->persistent, best_xcd = c->xcd;
    int best_tile = c->tile, best_tx = c->tile_x, best_na = c->node_acc;
    // node-accumulated deposit: mixed form, order 2
    const int nna = c->mixed && c->g.order == 2 ? 2 : 1;
    float best = 3.4e38f;
    for (int m = 0; m < nmap; ++m) {
    if (!det && c->interleave != maps[m]) {
        c->interleave = maps[m];
        apply_layout(c);
    }
    for (int tl = 0; tl < ntile; ++tl)
    for (int na = 0; na < nna; ++na)
    for (int tx = 1; tx <= 2; ++tx)
    for (int pm = 0; pm < npers; ++pm)
    for (int xo = 0; xo < 2; ++xo)
        for (int z = 0; z < 3; ++z) {
            if (sizes[z] > particle_max_threads(c->g, c->mixed)) continue;
            if (pm && xo) continue;  // the persistent form takes clusters from its work list
            if (tx == 2 && !(c->mixed && mixed_tile_x_ok(c->g, 2, pm != 0))) continue;
            if (na && !mixed_node_acc_ok(c->g, sizes[z], tx, c->interleave == 2, pm != 0)) continue;
            c->threads = sizes[z];
            c->node_acc = na;
            c->persistent = pm;
            c->xcd = xo;
            c->tile_x = tx;
            if (c->mixed) c->tile = 1 - tl;
            refresh_persistent(c);
            enqueue_particles(c, c->cur);  // warm-up
            SMI_HIP(hipEventRecord(c->ev[0], c->st));
            for (int r = 0; r < reps; ++r) enqueue_particles(c, c->cur);
            SMI_HIP(hipEventRecord(c->ev[1], c->st));
            SMI_HIP(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            SMI_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            ms /= reps;
            if (ms < best) {
                best = ms;
                best_map = c->interleave;
                best_threads = c->threads;
                best_pers = c->persistent;
                best_xcd = c->xcd;
                best_tile = c->tile;
                best_tx = c->tile_x;
                best_na = c->node_acc;
            }
        }
    }
    c->interleave = best_map;
    c->threads = best_threads;
    c->persistent = best_pers;
    c->xcd = best_xcd;
    c->tile = best_tile;
    c->tile_x = best_tx;
    c->node_acc = best_na;
    refresh_persistent(c);
    apply_layout(c);
    for (auto &s : c->sp) {
        launch_fill_i32(s.sb.counts, c->g.ncells, 0, c->st);
        launch_fill_i32((int32_t *)s.sb.stay, (c->g.ncells + 3) / 4, 0, c->st);
    }
    SMI_HIP(hipStreamSynchronize(c->st));
    if (best_ms) *best_ms = best;
    return SMI_OK;
    SMI_CATCH
}

namespace {

double reduce_partials(smi_context *c)
{
    std::vector<double> h(kReduceBlocks);
    SMI_HIP(hipMemcpyAsync(h.data(), c->partials, kReduceBlocks * sizeof(double), hipMemcpyDeviceToHost, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));
    double s = 0;
    for (double v : h) s += v;
    return s;
}

}  // namespace

int smi_scalars(smi_context *c, double *out)
{
    SMI_TRY
    check_ctx(c);
    if (!out) throw Error(SMI_ERR_INVALID, "null output");
    SMI_HIP(hipSetDevice(c->p.device));
    double cell_volume = 1.0;
    for (int d = 0; d < c->g.nd; ++d) cell_volume *= c->g.dx[d];
    const int fid[6] = {SMI_EX, SMI_EY, SMI_EZ, SMI_BXM, SMI_BYM, SMI_BZM};
    for (int f = 0; f < 6; ++f) {
        const double *one[1] = {c->F[fid[f]]};
        launch_sumsq(c->g, one, &fid[f], 1, c->partials, kReduceBlocks, c->st);
        out[f] = 0.5 * cell_volume * reduce_partials(c);
    }
    for (size_t k = 0; k < c->sp.size(); ++k) {
        SpeciesState &S = c->sp[k];
        double *o = out + 6 + 3 * k;
        o[0] = (double)S.n;
        if (c->mixed) launch_wsum(nullptr, (const float *)S.buf[c->cur][6], S.n, c->partials, kReduceBlocks, c->st);
        else launch_wsum((const double *)S.buf[c->cur][6], nullptr, S.n, c->partials, kReduceBlocks, c->st);
        o[1] = reduce_partials(c);
        if (c->mixed) launch_ukin_mixed(S.mview(c->cur), S.sb.perm, S.n, c->partials, kReduceBlocks, c->st);
        else launch_ukin(S.view(c->cur), S.sb.perm, S.n, c->partials, kReduceBlocks, c->st);
        o[2] = reduce_partials(c) * S.mass;
    }
    return SMI_OK;
    SMI_CATCH
}

int smi_energies(smi_context *c, double *out)
{
    SMI_TRY
    check_ctx(c);
    SMI_HIP(hipSetDevice(c->p.device));
    std::vector<double> h(kReduceBlocks);
    const int fid[6] = {SMI_EX, SMI_EY, SMI_EZ, SMI_BXM, SMI_BYM, SMI_BZM};
    const double *fl[6] = {c->F[SMI_EX], c->F[SMI_EY], c->F[SMI_EZ], c->F[SMI_BXM], c->F[SMI_BYM], c->F[SMI_BZM]};
    launch_sumsq(c->g, fl, fid, 6, c->partials, kReduceBlocks, c->st);
    SMI_HIP(hipMemcpyAsync(h.data(), c->partials, kReduceBlocks * sizeof(double), hipMemcpyDeviceToHost, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));
    double s = 0;
    for (double v : h) s += v;
    double cell_volume = 1.0;
    for (int d = 0; d < c->g.nd; ++d) cell_volume *= c->g.dx[d];
    out[0] = 0.5 * cell_volume * s;
    for (size_t k = 0; k < c->sp.size(); ++k) {
        SpeciesState &S = c->sp[k];
        if (c->mixed) launch_ukin_mixed(S.mview(c->cur), S.sb.perm, S.n, c->partials, kReduceBlocks, c->st);
        else launch_ukin(S.view(c->cur), S.sb.perm, S.n, c->partials, kReduceBlocks, c->st);
        SMI_HIP(hipMemcpyAsync(h.data(), c->partials, kReduceBlocks * sizeof(double), hipMemcpyDeviceToHost, c->st));
        SMI_HIP(hipStreamSynchronize(c->st));
        double e = 0;
        for (double v : h) e += v;
        out[1 + k] = e * S.mass;
    }
    return SMI_OK;
    SMI_CATCH
}

static void compute_rho(smi_context *c)
{
    launch_fill(c->rho, c->g.ncells, 0.0, c->st);
    for (auto &S : c->sp) {
        if (c->mixed) launch_rho_mixed(c->g, S.km, S.mview(c->cur), S.keys[c->cur], S.sb.perm, S.n, c->rho, c->st);
        else if (c->p.deposit == SMI_DEPOSIT_DETERMINISTIC)
            launch_rho_gather(c->g, S.k, S.view(c->cur), S.sb.perm, S.sb.cell_off, S.lay, c->rho, c->st);
        else launch_rho(c->g, S.k, S.view(c->cur), S.sb.perm, S.n, c->rho, c->st);
    }
}

int smi_rho(smi_context *c, double *gl)
{
    SMI_TRY
    check_ctx(c);
    SMI_HIP(hipSetDevice(c->p.device));
    compute_rho(c);
    SMI_HIP(hipMemcpyAsync(gl, c->rho, c->g.ncells * sizeof(double), hipMemcpyDeviceToHost, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));
    return SMI_OK;
    SMI_CATCH
}

int smi_gauss(smi_context *c, double *gl)
{
    SMI_TRY
    check_ctx(c);
    SMI_HIP(hipSetDevice(c->p.device));
    compute_rho(c);
    launch_gauss(c->g, c->fields(), c->rho, c->st);
    SMI_HIP(hipMemcpyAsync(gl, c->rho, c->g.ncells * sizeof(double), hipMemcpyDeviceToHost, c->st));
    SMI_HIP(hipStreamSynchronize(c->st));
    return SMI_OK;
    SMI_CATCH
}

}  // extern "C"
