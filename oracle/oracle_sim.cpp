// oracle_sim.cpp -- per-patch time step of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see oracle.h).
// The simulation is split into patches with ghosted field arrays exactly as
// the reference does (Patch::initStep3, Patch.cpp:136-172; ElectroMagn3D.cpp
// :88-180), and one time step runs the reference's sequence
// (Smilei.cpp:483-749):
//   per patch (OpenMP, VectorPatch.cpp:4777): restartRhoJ, then per species
//     gather -> Boris push -> boundary tagging -> Esirkepov deposit;
//   particle exchange to neighbour patches + cell sort;
//   J ghost sum in x, y, z passes (SyncVectorPatch::sumAllComponents);
//   saveMagneticFields, Maxwell-Ampere, Maxwell-Faraday, B exchange in
//   x, y, z (SyncVectorPatch::exchangeB), centerMagneticFields.
//
// Particle order inside a patch.  The reference's order of particles that
// share a cell is unspecified (its GPU sort is thrust::sort_by_key, its CPU
// sorts are not stable).  This oracle fixes one canonical order, which the
// HIP library reproduces: the particles arriving in patch D are listed by
// direction of motion m in {-1,0,1}^3 (index (mx+1)*9+(my+1)*3+(mz+1), the
// patch's own particles being direction 13), each group in its source
// patch's previous order, and that list is stable-sorted by cell key.
//
// Cell key (nvidiaParticles.cu:639-660, Cluster3D::CellKey): cell index
// ix = floor(x/dx) - Pcoord*patch (clamped to [0, patch-1] against rounding
// at the upper face), keys ordered by 4^3 clusters (4^2 in 2D), x-major.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// Staggering flags of the 12 stored components (Interpolator3D2Order.cpp:49-60).
const int kDual[12][3] = {
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},  // Ex Ey Ez
    {0, 1, 1}, {1, 0, 1}, {1, 1, 0},  // Bx By Bz
    {0, 1, 1}, {1, 0, 1}, {1, 1, 0},  // Bx_m By_m Bz_m
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},  // Jx Jy Jz
};
enum { EX, EY, EZ, BX, BY, BZ, BXM, BYM, BZM, JX, JY, JZ, NF };

struct Parts {
    std::vector<double> x, y, z, px, py, pz, w;
    std::vector<int32_t> key;
    size_t size() const { return w.size(); }
    void resize(size_t n)
    {
        x.resize(n); y.resize(n); z.resize(n);
        px.resize(n); py.resize(n); pz.resize(n);
        w.resize(n); key.resize(n);
    }
};

struct OPatch {
    int pc[3];
    int dom_begin[3];  // cell_starting_global_index (with ghosts)
    int noGC[3];       // cell_starting_global_index_noGC
    double min_local[3], max_local[3];
    std::vector<double> F[NF];
    std::vector<Parts> sp;
    std::vector<uint8_t> dir;  // scratch: direction of motion per particle
};

struct SpeciesDef {
    double mass, charge;
    double time_frozen = 0.;  // Species::time_frozen_
};

struct Sim {
    ora_params p;
    int nd;        // ndim
    int os;        // oversize
    int npatch[3]; // patches per dim
    int np[3];     // primal dims of a patch array
    int dims[NF][3];
    std::vector<OPatch> patches;
    std::vector<SpeciesDef> species;
    int64_t steps = 0;  // time steps done
    double timers[4] = {0, 0, 0, 0};

    // Species::dynamics pushes a species when time_dual > time_frozen
    // (Species.cpp:568, 832-866), time_dual = (steps + 1.5) dt during the
    // next step (Smilei.cpp:172, 487-488).
    bool moves(size_t s) const { return ((double)steps + 1.5) * p.timestep > species[s].time_frozen; }

    int pid(int a, int b, int c) const { return (a * npatch[1] + b) * npatch[2] + c; }
    size_t fsize(int c) const { return (size_t)dims[c][0] * dims[c][1] * dims[c][2]; }
};

inline int wrapi(int i, int n) { return ((i % n) + n) % n; }

// In-patch cell key (cluster-major, clusters of 4 cells per dimension).
int32_t cell_key(const Sim &S, const OPatch &P, const double *pos)
{
    int ic[3] = {0, 0, 0};
    for (int d = 0; d < S.nd; ++d) {
        const double inv = 1.0 / S.p.cell_length[d];
        int c = (int)std::floor(pos[d] * inv) - P.noGC[d];
        if (c < 0) c = 0;
        if (c > S.p.patch_size[d] - 1) c = S.p.patch_size[d] - 1;
        ic[d] = c;
    }
    if (S.nd == 3) {
        const int ncy = S.p.patch_size[1] / 4, ncz = S.p.patch_size[2] / 4;
        const int cl = ((ic[0] / 4) * ncy + ic[1] / 4) * ncz + ic[2] / 4;
        return cl * 64 + 16 * (ic[0] % 4) + 4 * (ic[1] % 4) + ic[2] % 4;
    }
    const int ncy = S.p.patch_size[1] / 4;
    const int cl = (ic[0] / 4) * ncy + ic[1] / 4;
    return cl * 16 + 4 * (ic[0] % 4) + ic[1] % 4;
}

int cells_per_patch(const Sim &S)
{
    int c = 1;
    for (int d = 0; d < S.nd; ++d) c *= S.p.patch_size[d];
    return c;
}

// Stable counting sort of a patch's particle list by cell key.
void sort_patch(const Sim &S, Parts &pt)
{
    const size_t n = pt.size();
    const int nk = cells_per_patch(S);
    std::vector<int64_t> off(nk + 1, 0);
    for (size_t i = 0; i < n; ++i) off[pt.key[i] + 1]++;
    for (int k = 0; k < nk; ++k) off[k + 1] += off[k];
    Parts o;
    o.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const int64_t j = off[pt.key[i]]++;
        o.x[j] = pt.x[i]; o.y[j] = pt.y[i]; o.z[j] = pt.z[i];
        o.px[j] = pt.px[i]; o.py[j] = pt.py[i]; o.pz[j] = pt.pz[i];
        o.w[j] = pt.w[i]; o.key[j] = pt.key[i];
    }
    pt = std::move(o);
}

// Patch coordinate along d holding position v: min_local <= v < max_local,
// with min_local = Pcoord*(patch*dx) (Patch.cpp:149-150).
int locate(const Sim &S, int d, double v)
{
    const double L = (double)S.p.patch_size[d] * S.p.cell_length[d];
    int c = (int)std::floor(v / L);
    if (c < 0) c = 0;
    if (c > S.npatch[d] - 1) c = S.npatch[d] - 1;
    while (c > 0 && v < (double)c * L) --c;
    while (c < S.npatch[d] - 1 && v >= (double)(c + 1) * L) ++c;
    return c;
}

// ---------------------------------------------------------------- dynamics
void patch_dynamics(Sim &S, OPatch &P)
{
    for (int c = JX; c <= JZ; ++c) std::fill(P.F[c].begin(), P.F[c].end(), 0.0);  // restartRhoJ
    const double dinv[3] = {1.0 / S.p.cell_length[0], 1.0 / S.p.cell_length[1], 1.0 / S.p.cell_length[2]};
    double cell_volume = 1.0;
    for (int d = 0; d < S.nd; ++d) cell_volume *= S.p.cell_length[d];
    const double icv = 1. / cell_volume;
    std::vector<double> Ep, Bp, invgf, dold;
    std::vector<int32_t> iold;
    for (size_t s = 0; s < S.species.size(); ++s) {
        Parts &pt = P.sp[s];
        const int64_t n = (int64_t)pt.size();
        if (n == 0 || !S.moves(s)) continue;
        Ep.resize(3 * n); Bp.resize(3 * n); invgf.resize(n);
        dold.resize(3 * n); iold.resize(3 * n);
        if (S.nd == 3) {
            ora_interp3d(S.p.order, P.F[EX].data(), P.F[EY].data(), P.F[EZ].data(),
                         P.F[BXM].data(), P.F[BYM].data(), P.F[BZM].data(), S.np, P.dom_begin, dinv,
                         n, pt.x.data(), pt.y.data(), pt.z.data(), Ep.data(), Bp.data(), iold.data(), dold.data());
        } else {
            ora_interp2d(S.p.order, P.F[EX].data(), P.F[EY].data(), P.F[EZ].data(),
                         P.F[BXM].data(), P.F[BYM].data(), P.F[BZM].data(), S.np, P.dom_begin, dinv,
                         n, pt.x.data(), pt.y.data(), Ep.data(), Bp.data(), iold.data(), dold.data());
        }
        ora_push_boris(S.nd, S.species[s].charge, S.species[s].mass, S.p.timestep, n,
                       pt.x.data(), pt.y.data(), pt.z.data(), pt.px.data(), pt.py.data(), pt.pz.data(),
                       Ep.data(), Bp.data(), invgf.data());
        if (S.nd == 3) {
            ora_proj3d(S.p.order, P.F[JX].data(), P.F[JY].data(), P.F[JZ].data(), S.np, P.dom_begin,
                       S.p.cell_length, S.p.timestep, icv, S.species[s].charge, n,
                       pt.x.data(), pt.y.data(), pt.z.data(), pt.w.data(), iold.data(), dold.data());
        } else {
            ora_proj2d(S.p.order, P.F[JX].data(), P.F[JY].data(), P.F[JZ].data(), S.np, P.dom_begin,
                       S.p.cell_length, S.p.timestep, icv, S.species[s].charge, n,
                       pt.x.data(), pt.y.data(), pt.pz.data(), pt.w.data(), invgf.data(),
                       iold.data(), dold.data());
        }
    }
}

// ------------------------------------------------ exchange + sort (per species)
// Boundary tagging of PartBoundCond::apply / internal_inf|sup
// (BoundaryConditionType.cpp:15-56) decides the direction of motion per
// dimension; the periodic wrap of Patch::prepareParticles (Patch.cpp:
// 624-668) is applied on arrival; the particle then belongs to the patch
// reached by the x, y, z passes with corner forwarding (Patch.cpp:727).
void exchange_and_sort(Sim &S, int s)
{
    const int npt = (int)S.patches.size();
    double xmax[3];
    for (int d = 0; d < 3; ++d) xmax[d] = S.p.cell_length[d] * (double)S.p.n_cells[d];
#pragma omp parallel for schedule(dynamic)
    for (int ip = 0; ip < npt; ++ip) {
        OPatch &P = S.patches[ip];
        Parts &pt = P.sp[s];
        P.dir.resize(pt.size());
        for (size_t i = 0; i < pt.size(); ++i) {
            const double pos[3] = {pt.x[i], pt.y[i], pt.z[i]};
            int m[3] = {0, 0, 0};
            for (int d = 0; d < S.nd; ++d) {
                if (pos[d] < P.min_local[d]) m[d] = -1;
                else if (pos[d] >= P.max_local[d]) m[d] = 1;
            }
            P.dir[i] = (uint8_t)((m[0] + 1) * 9 + (m[1] + 1) * 3 + (m[2] + 1));
        }
    }
    std::vector<Parts> incoming(npt);
#pragma omp parallel for schedule(dynamic)
    for (int ip = 0; ip < npt; ++ip) {
        OPatch &D = S.patches[ip];
        Parts &in = incoming[ip];
        for (int dir = 0; dir < 27; ++dir) {
            const int m[3] = {dir / 9 - 1, (dir / 3) % 3 - 1, dir % 3 - 1};
            if (S.nd == 2 && m[2] != 0) continue;
            int sc[3];
            for (int d = 0; d < 3; ++d) sc[d] = wrapi(D.pc[d] - m[d], S.npatch[d]);
            const OPatch &Src = S.patches[S.pid(sc[0], sc[1], sc[2])];
            const Parts &sp = Src.sp[s];
            for (size_t i = 0; i < sp.size(); ++i) {
                if (Src.dir[i] != dir) continue;
                double pos[3] = {sp.x[i], sp.y[i], sp.z[i]};
                for (int d = 0; d < S.nd; ++d) {
                    if (m[d] == -1 && Src.pc[d] == 0 && pos[d] < 0.) pos[d] += xmax[d];
                    if (m[d] == 1 && Src.pc[d] == S.npatch[d] - 1 && pos[d] >= xmax[d]) pos[d] -= xmax[d];
                }
                in.x.push_back(pos[0]); in.y.push_back(pos[1]); in.z.push_back(pos[2]);
                in.px.push_back(sp.px[i]); in.py.push_back(sp.py[i]); in.pz.push_back(sp.pz[i]);
                in.w.push_back(sp.w[i]);
                in.key.push_back(cell_key(S, D, pos));
            }
        }
        sort_patch(S, in);
    }
    for (int ip = 0; ip < npt; ++ip) S.patches[ip].sp[s] = std::move(incoming[ip]);
}

// --------------------------------------------------------- J ghost sum
// SyncVectorPatch::sumAllComponents (SyncVectorPatch.cpp:203-~630): along
// dimension d the low neighbour's planes [n, n+gsp) and this patch's planes
// [0, gsp), gsp = 1 + 2*os + dual_d, are replaced by (low + this).
// arr(patch, component) gives a patch's array of J component JX+c.
template <class Arr>
void sum_ghosts_on(const Sim &S, Arr arr)
{
    const int npt = (int)S.patches.size();
    for (int d = 0; d < S.nd; ++d) {
        for (int c = JX; c <= JZ; ++c) {
            const int *dm = S.dims[c];
            const int gsp = 1 + 2 * S.os + kDual[c][d];
            const int n = S.p.patch_size[d];
#pragma omp parallel for schedule(static)
            for (int ip = 0; ip < npt; ++ip) {
                const OPatch &P = S.patches[ip];
                int lc[3] = {P.pc[0], P.pc[1], P.pc[2]};
                lc[d] = wrapi(lc[d] - 1, S.npatch[d]);
                double *pt = arr(ip, c);
                double *lo = arr(S.pid(lc[0], lc[1], lc[2]), c);
                int r0[3] = {0, 0, 0}, r1[3] = {dm[0], dm[1], dm[2]};
                r1[d] = gsp;
                for (int i = r0[0]; i < r1[0]; ++i)
                    for (int j = r0[1]; j < r1[1]; ++j)
                        for (int k = r0[2]; k < r1[2]; ++k) {
                            int li[3] = {i, j, k};
                            li[d] += n;
                            const size_t a = ((size_t)li[0] * dm[1] + li[1]) * dm[2] + li[2];
                            const size_t b = ((size_t)i * dm[1] + j) * dm[2] + k;
                            lo[a] += pt[b];
                            pt[b] = lo[a];
                        }
            }
        }
    }
}

void sum_ghosts(Sim &S)
{
    sum_ghosts_on(S, [&](int ip, int c) { return S.patches[ip].F[c].data(); });
}

// --------------------------------------------------------- B exchange
// SyncVectorPatch::exchangeB -> exchangeAllComponentsAlong{X,Y,Z}
// (SyncVectorPatch.cpp:667, 1441-1830): for the two B components dual in d,
// this patch's planes [0, os) <- low neighbour's [n, n+os), and the low
// neighbour's [n+gsp, n+gsp+os) <- this patch's [gsp, gsp+os), gsp = os+2.
// arr(patch, component) gives a patch's array of B component BX+c.
template <class Arr>
void exchange_B_on(const Sim &S, Arr arr)
{
    const int npt = (int)S.patches.size();
    for (int d = 0; d < S.nd; ++d) {
        for (int c = BX; c <= BZ; ++c) {
            if (!kDual[c][d]) continue;
            const int *dm = S.dims[c];
            const int os = S.os, gsp = os + 2, n = S.p.patch_size[d];
#pragma omp parallel for schedule(static)
            for (int ip = 0; ip < npt; ++ip) {
                const OPatch &P = S.patches[ip];
                int lc[3] = {P.pc[0], P.pc[1], P.pc[2]};
                lc[d] = wrapi(lc[d] - 1, S.npatch[d]);
                double *pt = arr(ip, c);
                double *lo = arr(S.pid(lc[0], lc[1], lc[2]), c);
                int r1[3] = {dm[0], dm[1], dm[2]};
                r1[d] = os;
                for (int i = 0; i < r1[0]; ++i)
                    for (int j = 0; j < r1[1]; ++j)
                        for (int k = 0; k < r1[2]; ++k) {
                            int a[3] = {i, j, k}, b[3] = {i, j, k};
                            a[d] += n;  // low neighbour [n, n+os)  -> this [0, os)
                            pt[((size_t)b[0] * dm[1] + b[1]) * dm[2] + b[2]] =
                                lo[((size_t)a[0] * dm[1] + a[1]) * dm[2] + a[2]];
                            int e[3] = {i, j, k}, f[3] = {i, j, k};
                            e[d] += n + gsp;  // low [n+gsp, ..) <- this [gsp, ..)
                            f[d] += gsp;
                            lo[((size_t)e[0] * dm[1] + e[1]) * dm[2] + e[2]] =
                                pt[((size_t)f[0] * dm[1] + f[1]) * dm[2] + f[2]];
                        }
            }
        }
    }
}

void exchange_B(Sim &S)
{
    exchange_B_on(S, [&](int ip, int c) { return S.patches[ip].F[c].data(); });
}

void maxwell(Sim &S)
{
    const int npt = (int)S.patches.size();
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < npt; ++ip) {
        OPatch &P = S.patches[ip];
        P.F[BXM] = P.F[BX];  // saveMagneticFields (ElectroMagn3D.cpp:1048)
        P.F[BYM] = P.F[BY];
        P.F[BZM] = P.F[BZ];
        if (S.nd == 3)
            ora_yee3d_ma(P.F[EX].data(), P.F[EY].data(), P.F[EZ].data(), P.F[BX].data(), P.F[BY].data(),
                         P.F[BZ].data(), P.F[JX].data(), P.F[JY].data(), P.F[JZ].data(), S.np, S.p.timestep,
                         S.p.cell_length);
        else
            ora_yee2d_ma(P.F[EX].data(), P.F[EY].data(), P.F[EZ].data(), P.F[BX].data(), P.F[BY].data(),
                         P.F[BZ].data(), P.F[JX].data(), P.F[JY].data(), P.F[JZ].data(), S.np, S.p.timestep,
                         S.p.cell_length);
    }
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < npt; ++ip) {
        OPatch &P = S.patches[ip];
        if (S.nd == 3)
            ora_yee3d_mf(P.F[BX].data(), P.F[BY].data(), P.F[BZ].data(), P.F[EX].data(), P.F[EY].data(),
                         P.F[EZ].data(), S.np, S.p.timestep, S.p.cell_length);
        else
            ora_yee2d_mf(P.F[BX].data(), P.F[BY].data(), P.F[BZ].data(), P.F[EX].data(), P.F[EY].data(),
                         P.F[EZ].data(), S.np, S.p.timestep, S.p.cell_length);
    }
    exchange_B(S);
    // centerMagneticFields (ElectroMagn3D.cpp:1191): B_m = (B + B_m) * 0.5
#pragma omp parallel for schedule(static)
    for (int ip = 0; ip < npt; ++ip) {
        OPatch &P = S.patches[ip];
        for (int c = 0; c < 3; ++c) {
            double *b = P.F[BX + c].data(), *bm = P.F[BXM + c].data();
            const size_t sz = P.F[BX + c].size();
            for (size_t e = 0; e < sz; ++e) bm[e] = (b[e] + bm[e]) * 0.5;
        }
    }
}

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void step(Sim &S)
{
    const int npt = (int)S.patches.size();
    double t0 = now();
#pragma omp parallel for schedule(dynamic)
    for (int ip = 0; ip < npt; ++ip) patch_dynamics(S, S.patches[ip]);
    double t1 = now();
    for (size_t s = 0; s < S.species.size(); ++s)
        if (S.moves(s)) exchange_and_sort(S, (int)s);
    double t2 = now();
    sum_ghosts(S);
    double t3 = now();
    maxwell(S);
    double t4 = now();
    S.timers[0] += t1 - t0;
    S.timers[1] += t2 - t1;
    S.timers[2] += t3 - t2;
    S.timers[3] += t4 - t3;
    S.steps += 1;
}

// Global index of local point l of patch P along d (primal and dual share
// the origin dom_begin; periodic).
inline int gidx(const Sim &S, const OPatch &P, int d, int l)
{
    return wrapi(P.dom_begin[d] + l, S.p.n_cells[d]);
}

}  // namespace

extern "C" {

void *ora_create(const ora_params *p)
{
    if (!p || (p->ndim != 2 && p->ndim != 3) || (p->order != 2 && p->order != 4)) return nullptr;
    Sim *S = new Sim();
    S->p = *p;
    S->nd = p->ndim;
    S->os = p->order;
    for (int d = 0; d < 3; ++d) {
        if (d >= S->nd) {
            S->p.n_cells[d] = 1;
            S->p.patch_size[d] = 1;
            if (S->p.cell_length[d] <= 0) S->p.cell_length[d] = 1.0;
        }
        if (S->p.n_cells[d] % S->p.patch_size[d] != 0) { delete S; return nullptr; }
        if (d < S->nd && S->p.patch_size[d] % 4 != 0) { delete S; return nullptr; }
        S->npatch[d] = S->p.n_cells[d] / S->p.patch_size[d];
        S->np[d] = d < S->nd ? S->p.patch_size[d] + 2 * S->os + 1 : 1;
    }
    for (int c = 0; c < NF; ++c)
        for (int d = 0; d < 3; ++d) S->dims[c][d] = S->np[d] + (d < S->nd ? kDual[c][d] : 0);
    const int npt = S->npatch[0] * S->npatch[1] * S->npatch[2];
    S->patches.resize(npt);
    for (int a = 0; a < S->npatch[0]; ++a)
        for (int b = 0; b < S->npatch[1]; ++b)
            for (int c = 0; c < S->npatch[2]; ++c) {
                OPatch &P = S->patches[S->pid(a, b, c)];
                P.pc[0] = a; P.pc[1] = b; P.pc[2] = c;
                for (int d = 0; d < 3; ++d) {
                    const double L = (double)S->p.patch_size[d] * S->p.cell_length[d];
                    P.min_local[d] = (double)P.pc[d] * L;
                    P.max_local[d] = (double)(P.pc[d] + 1) * L;
                    P.noGC[d] = P.pc[d] * S->p.patch_size[d];
                    P.dom_begin[d] = P.noGC[d] - (d < S->nd ? S->os : 0);
                }
                for (int f = 0; f < NF; ++f) P.F[f].assign(S->fsize(f), 0.0);
            }
    return S;
}

void ora_destroy(void *h) { delete (Sim *)h; }

void ora_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int ora_set_time_frozen(void *h, int s, double time_frozen)
{
    Sim *S = (Sim *)h;
    if (s < 0 || s >= (int)S->species.size()) return -1;
    S->species[s].time_frozen = time_frozen;
    return 0;
}

int ora_add_species(void *h, double mass, double charge)
{
    Sim *S = (Sim *)h;
    S->species.push_back({mass, charge});
    for (auto &P : S->patches) P.sp.emplace_back();
    return (int)S->species.size() - 1;
}

int ora_set_particles(void *h, int s, int64_t n, const double *x, const double *y, const double *z,
                      const double *px, const double *py, const double *pz, const double *w)
{
    Sim *S = (Sim *)h;
    if (s < 0 || s >= (int)S->species.size()) return -1;
    for (auto &P : S->patches) P.sp[s] = Parts();
    for (int64_t i = 0; i < n; ++i) {
        double pos[3] = {x[i], S->nd > 1 ? y[i] : 0., S->nd > 2 ? z[i] : 0.};
        int c[3] = {0, 0, 0};
        for (int d = 0; d < S->nd; ++d) {
            const double xm = S->p.cell_length[d] * (double)S->p.n_cells[d];
            if (!(pos[d] >= 0. && pos[d] < xm)) return -2;
            c[d] = locate(*S, d, pos[d]);
        }
        OPatch &P = S->patches[S->pid(c[0], c[1], c[2])];
        Parts &pt = P.sp[s];
        pt.x.push_back(pos[0]); pt.y.push_back(pos[1]); pt.z.push_back(pos[2]);
        pt.px.push_back(px[i]); pt.py.push_back(py[i]); pt.pz.push_back(pz[i]);
        pt.w.push_back(w[i]);
        pt.key.push_back(cell_key(*S, P, pos));
    }
    for (auto &P : S->patches) sort_patch(*S, P.sp[s]);
    return 0;
}

int64_t ora_num_particles(void *h, int s)
{
    Sim *S = (Sim *)h;
    int64_t n = 0;
    for (auto &P : S->patches) n += (int64_t)P.sp[s].size();
    return n;
}

int ora_get_particles(void *h, int s, double *x, double *y, double *z, double *px, double *py,
                      double *pz, double *w, int32_t *patch, int32_t *key)
{
    Sim *S = (Sim *)h;
    size_t o = 0;
    for (size_t ip = 0; ip < S->patches.size(); ++ip) {
        const Parts &pt = S->patches[ip].sp[s];
        for (size_t i = 0; i < pt.size(); ++i, ++o) {
            if (x) x[o] = pt.x[i];
            if (y) y[o] = pt.y[i];
            if (z) z[o] = pt.z[i];
            if (px) px[o] = pt.px[i];
            if (py) py[o] = pt.py[i];
            if (pz) pz[o] = pt.pz[i];
            if (w) w[o] = pt.w[i];
            if (patch) patch[o] = (int32_t)ip;
            if (key) key[o] = pt.key[i];
        }
    }
    return 0;
}

int ora_patch_counts(void *h, int s, int32_t *counts)
{
    Sim *S = (Sim *)h;
    for (size_t ip = 0; ip < S->patches.size(); ++ip) counts[ip] = (int32_t)S->patches[ip].sp[s].size();
    return 0;
}

int ora_set_field(void *h, int fid, const double *g)
{
    Sim *S = (Sim *)h;
    if (fid < 0 || fid >= NF) return -1;
    const int N1 = S->p.n_cells[1], N2 = S->p.n_cells[2];
    for (auto &P : S->patches) {
        const int *dm = S->dims[fid];
        for (int i = 0; i < dm[0]; ++i)
            for (int j = 0; j < dm[1]; ++j)
                for (int k = 0; k < dm[2]; ++k) {
                    const size_t gi = ((size_t)gidx(*S, P, 0, i) * N1 + gidx(*S, P, 1, j)) * N2 + gidx(*S, P, 2, k);
                    P.F[fid][((size_t)i * dm[1] + j) * dm[2] + k] = g[gi];
                }
    }
    return 0;
}

int ora_get_field(void *h, int fid, double *g)
{
    Sim *S = (Sim *)h;
    if (fid < 0 || fid >= NF) return -1;
    const int N1 = S->p.n_cells[1], N2 = S->p.n_cells[2];
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = d < S->nd ? S->os : 0;
        hi[d] = lo[d] + S->p.patch_size[d];
    }
    for (auto &P : S->patches) {
        const int *dm = S->dims[fid];
        for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
                for (int k = lo[2]; k < hi[2]; ++k) {
                    const size_t gi = ((size_t)gidx(*S, P, 0, i) * N1 + gidx(*S, P, 1, j)) * N2 + gidx(*S, P, 2, k);
                    g[gi] = P.F[fid][((size_t)i * dm[1] + j) * dm[2] + k];
                }
    }
    return 0;
}

int ora_step(void *h, int nsteps)
{
    Sim *S = (Sim *)h;
    for (int t = 0; t < 4; ++t) S->timers[t] = 0;
    for (int it = 0; it < nsteps; ++it) step(*S);
    return 0;
}

int ora_timers(void *h, double *out)
{
    Sim *S = (Sim *)h;
    for (int t = 0; t < 4; ++t) out[t] = S->timers[t];
    return 0;
}

}  // extern "C"

namespace {

// Local index range [lo, lo+n) of a patch array along axis d that
// Field3D::norm2 sums (istart / bufsize of ElectroMagn3D.cpp:185-226): the
// "non duplicated elements", which with periodic boundaries still include
// the shared face point of the first patch (primal) and one face point at
// each end of the domain (dual), i.e. N+1 primal and N+2 dual points per
// axis over all patches.
void norm2_range(const Sim &S, const OPatch &P, int d, int dual, int &lo, int &n)
{
    if (d >= S.nd) {
        lo = 0;
        n = 1;
        return;
    }
    lo = S.os + (P.pc[d] != 0 ? 1 : 0);
    n = S.p.patch_size[d] + 1 + dual;
    if (S.npatch[d] != 1) {
        if (!dual && P.pc[d] != 0) {
            n -= 1;
        } else if (dual) {
            n -= 1;
            if (P.pc[d] != 0 && P.pc[d] != S.npatch[d] - 1) n -= 1;
        }
    }
}

// Field3D::norm2 (Field3D.cpp:230-250) of component f summed over patches.
double field_norm2(const Sim &S, int f)
{
    double a = 0.;
    for (const auto &P : S.patches) {
        int lo[3], n[3];
        for (int d = 0; d < 3; ++d) norm2_range(S, P, d, kDual[f][d], lo[d], n[d]);
        const int *dm = S.dims[f];
        for (int i = lo[0]; i < lo[0] + n[0]; ++i)
            for (int j = lo[1]; j < lo[1] + n[1]; ++j)
                for (int k = lo[2]; k < lo[2] + n[2]; ++k) {
                    const double v = P.F[f][((size_t)i * dm[1] + j) * dm[2] + k];
                    a += v * v;
                }
    }
    return a;
}

}  // namespace

extern "C" {

// Per-field electromagnetic energies of DiagnosticScalar::compute
// (DiagnosticScalar.cpp:660-690): Uelm_f = 0.5 * V_cell * norm2(f) for
// Ex, Ey, Ez, Bx_m, By_m, Bz_m.
int ora_field_energies(void *h, double *out)
{
    Sim *S = (Sim *)h;
    double cell_volume = 1.0;
    for (int d = 0; d < S->nd; ++d) cell_volume *= S->p.cell_length[d];
    const int fl[6] = {EX, EY, EZ, BXM, BYM, BZM};
    for (int q = 0; q < 6; ++q) out[q] = 0.5 * cell_volume * field_norm2(*S, fl[q]);
    return 0;
}

// DiagnosticScalar (DiagnosticScalar.cpp:494-522, 660-690): Ukin_s = m_s *
// sum w (gamma - 1); Uelm = sum of the six per-field energies
// (ElectroMagn::computeEnergy, ElectroMagn.cpp:465-477).
int ora_energies(void *h, double *out)
{
    Sim *S = (Sim *)h;
    double fe[6];
    ora_field_energies(h, fe);
    out[0] = fe[0] + fe[1] + fe[2] + fe[3] + fe[4] + fe[5];
    for (size_t s = 0; s < S->species.size(); ++s) {
        double e = 0.;
        for (auto &P : S->patches) {
            const Parts &pt = P.sp[s];
            for (size_t i = 0; i < pt.size(); ++i) {
                const double gm = std::sqrt(1 + pt.px[i] * pt.px[i] + pt.py[i] * pt.py[i] + pt.pz[i] * pt.pz[i]);
                e += pt.w[i] * (gm - 1.0);
            }
        }
        out[1 + s] = e * S->species[s].mass;
    }
    return 0;
}

int ora_rho(void *h, double *g)
{
    Sim *S = (Sim *)h;
    const int N[3] = {S->p.n_cells[0], S->p.n_cells[1], S->p.n_cells[2]};
    const size_t NN = (size_t)N[0] * N[1] * N[2];
    std::fill(g, g + NN, 0.0);
    double cell_volume = 1.0;
    for (int d = 0; d < S->nd; ++d) cell_volume *= S->p.cell_length[d];
    const double icv = 1. / cell_volume;
    const int hh = S->p.order / 2, L = 2 * hh + 3;
    for (size_t s = 0; s < S->species.size(); ++s)
        for (auto &P : S->patches) {
            const Parts &pt = P.sp[s];
            for (size_t i = 0; i < pt.size(); ++i) {
                const double pos[3] = {pt.x[i], pt.y[i], pt.z[i]};
                double Sw[3][7];
                int o[3] = {0, 0, 0};
                for (int d = 0; d < 3; ++d) {
                    for (int a = 0; a < L; ++a) Sw[d][a] = 0.;
                    if (d >= S->nd) { Sw[d][0] = 1.0; continue; }
                    const double pn = pos[d] * (1.0 / S->p.cell_length[d]);
                    const int ipn = (int)std::round(pn);
                    const double dd = pn - (double)ipn;
                    if (S->p.order == 2) {
                        const double d2 = dd * dd;
                        Sw[d][1] = 0.5 * (d2 - dd + 0.25);
                        Sw[d][2] = 0.75 - d2;
                        Sw[d][3] = 0.5 * (d2 + dd + 0.25);
                    } else {
                        const double d2 = dd * dd, d3 = d2 * dd, d4 = d3 * dd;
                        Sw[d][1] = 1.0 / 384.0 - 1.0 / 48.0 * dd + 1.0 / 16.0 * d2 - 1.0 / 12.0 * d3 + 1.0 / 24.0 * d4;
                        Sw[d][2] = 19.0 / 96.0 - 11.0 / 24.0 * dd + 0.25 * d2 + 1.0 / 6.0 * d3 - 1.0 / 6.0 * d4;
                        Sw[d][3] = 115.0 / 192.0 - 5.0 / 8.0 * d2 + 0.25 * d4;
                        Sw[d][4] = 19.0 / 96.0 + 11.0 / 24.0 * dd + 0.25 * d2 - 1.0 / 6.0 * d3 - 1.0 / 6.0 * d4;
                        Sw[d][5] = 1.0 / 384.0 + 1.0 / 48.0 * dd + 1.0 / 16.0 * d2 + 1.0 / 12.0 * d3 + 1.0 / 24.0 * d4;
                    }
                    o[d] = ipn - (hh + 1);
                }
                const double cw = icv * S->species[s].charge * pt.w[i];
                const int Lx = L, Ly = S->nd > 1 ? L : 1, Lz = S->nd > 2 ? L : 1;
                for (int a = 0; a < Lx; ++a)
                    for (int b = 0; b < Ly; ++b)
                        for (int c = 0; c < Lz; ++c) {
                            const double v = cw * Sw[0][a] * Sw[1][b] * Sw[2][c];
                            if (v == 0.) continue;
                            const size_t gi = ((size_t)wrapi(o[0] + a, N[0]) * N[1] + wrapi(o[1] + b, N[1])) * N[2] +
                                              wrapi(o[2] + c, N[2]);
                            g[gi] += v;
                        }
            }
        }
    return 0;
}

// Offsets of the three components inside one patch's block of a
// patch-major flat array (components first..first+2 of the patch layout).
static void flat_offsets(const Sim &S, int first, size_t off[4])
{
    off[0] = 0;
    for (int c = 0; c < 3; ++c) off[c + 1] = off[c] + S.fsize(first + c);
}

// J ghost sum on patch-major flat arrays: patch p's block at p * (|Jx| +
// |Jy| + |Jz|), holding Jx, Jy, Jz in the per-patch ghosted layout.
int ora_sum_j_ghosts(const ora_params *p, double *J)
{
    Sim *S = (Sim *)ora_create(p);
    if (!S) return -1;
    size_t off[4];
    flat_offsets(*S, JX, off);
    sum_ghosts_on(*S, [&](int ip, int c) { return J + (size_t)ip * off[3] + off[c - JX]; });
    ora_destroy(S);
    return 0;
}

// B exchange on patch-major flat arrays [patch][Bx|By|Bz] (ghosted layout).
int ora_exchange_b(const ora_params *p, double *B)
{
    Sim *S = (Sim *)ora_create(p);
    if (!S) return -1;
    size_t off[4];
    flat_offsets(*S, BX, off);
    exchange_B_on(*S, [&](int ip, int c) { return B + (size_t)ip * off[3] + off[c - BX]; });
    ora_destroy(S);
    return 0;
}

// div(E) - rho on primal nodes; dual index g+1 sits at (g+0.5)*dx.
int ora_gauss(void *h, double *g)
{
    Sim *S = (Sim *)h;
    const int N[3] = {S->p.n_cells[0], S->p.n_cells[1], S->p.n_cells[2]};
    const size_t NN = (size_t)N[0] * N[1] * N[2];
    std::vector<double> E[3];
    for (int c = 0; c < 3; ++c) {
        E[c].resize(NN);
        ora_get_field(h, EX + c, E[c].data());
    }
    ora_rho(h, g);
    auto at = [&](int c, int i, int j, int k) {
        return E[c][((size_t)wrapi(i, N[0]) * N[1] + wrapi(j, N[1])) * N[2] + wrapi(k, N[2])];
    };
    for (int i = 0; i < N[0]; ++i)
        for (int j = 0; j < N[1]; ++j)
            for (int k = 0; k < N[2]; ++k) {
                double div = (at(0, i + 1, j, k) - at(0, i, j, k)) / S->p.cell_length[0];
                if (S->nd > 1) div += (at(1, i, j + 1, k) - at(1, i, j, k)) / S->p.cell_length[1];
                if (S->nd > 2) div += (at(2, i, j, k + 1) - at(2, i, j, k)) / S->p.cell_length[2];
                const size_t gi = ((size_t)i * N[1] + j) * N[2] + k;
                g[gi] = div - g[gi];
            }
    return 0;
}

}  // extern "C"
