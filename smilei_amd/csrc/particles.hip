// particles.hip -- the particle half of the PIC step on gfx950, FP64 form.
//
// k_particles<ND,ORDER,DEPOSIT>: one workgroup per 4-cell cluster of every
// patch (the reference's GPU cluster, Params.h:403-418).  The cluster's
// particles are contiguous in the cell-sorted list.  The workgroup
//   1. stages the six gathered field components of the cluster plus its
//      stencil halo in LDS,
//   2. per particle: gathers E and B_m (Interpolator{2D2,3D2,3D4}Order),
//      Boris-pushes (PusherBoris), tags the direction of motion against the
//      patch limits (internal_inf/sup), wraps periodic leavers
//      (Patch::prepareParticles), computes the destination cell key
//      (Cluster3D::CellKey) and, with DEPOSIT, adds the Esirkepov currents
//      (Projector{2D2,3D2,3D4}Order::currents) into an LDS tile with LDS
//      FP64 atomics,
//   3. flushes the J tile into the global arrays with FP64 atomics.
// Gather, push and the per-particle current contributions use the
// reference's association order (the file is built with -ffp-contract=off),
// so only the order of the J summation differs from the CPU operators.
//
// k_det_deposit: the same current contributions added in the reference's
// order into per-patch ghosted arrays, one wavefront per patch walking the
// patch's particles in sequence; then the J ghost sum in x, y, z passes
// (SyncVectorPatch::sumAllComponents) and a copy of the patch interiors.
#include "kernels.h"
#include "pic_device.h"

namespace smi {

namespace {

using namespace dev;

// Esirkepov currents of one 3D particle into the cluster's LDS tile: old
// delta dold, new position in cell units pnn, old node ip; NT as
// esirkepov3d_win.
template <int ORDER, int NT>
__device__ inline void deposit3_tile(double *sJ, const Geom &g, const int *G0, const double *dold, const double *pnn,
                                     const int *ip, double cw)
{
    using TL = Tile<ORDER>;
    constexpr int H = TL::H, JT = TL::JT;
    // J-tile strides: order-2 rows padded by one word (fp_cluster's layout)
    constexpr int JSY = ORDER == 2 ? JT + 1 : JT, JSX = JT * JSY, JSC = JT * JSX;
    double S0[3][ORDER + 2], DS[3][ORDER + 2];
    int w0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) esirkepov_window<ORDER>(dold[d], pnn[d], ip[d], S0[d], DS[d], w0[d]);
    const double cr[3] = {cw * (g.dx[0] / g.dt), cw * (g.dx[1] / g.dt), cw * (g.dx[2] / g.dt)};
    // tile coords of stencil index 0 along each axis
    const int o0 = ip[0] - (H + 1) - (G0[0] - TL::JO);
    const int o1 = ip[1] - (H + 1) - (G0[1] - TL::JO);
    const int o2 = ip[2] - (H + 1) - (G0[2] - TL::JO);
    double *sJb = sJ + o0 * JSX + o1 * JSY + o2;
    esirkepov3d_win<ORDER, NT>(S0, DS, w0, cr, [&](int c, int a, int b, int k, double v) {
        atomicAdd(&sJb[c * JSC + a * JSX + b * JSY + k], v);
    });
}

// ---------------------------------------------------------------------------
// Fused particle kernel.
// ---------------------------------------------------------------------------
// One workgroup runs the particles of every species of the launch
// (specs[0..ns)) of cluster `cluster` against one field tile (and, with
// DEPOSIT, one J tile).  Uniform across the workgroup; returns at once for
// an empty cluster.
// Inlined into its kernels: left to itself the compiler calls it (order 4),
// and the call's stack frame lives in scratch memory.
template <int ND, int ORDER, bool DEPOSIT>
__device__ __attribute__((always_inline)) inline void fp_cluster(const Geom &g, const FpSpecies *__restrict__ specs,
                                                                 int ns, const GatherFields &F, const JPtrs &J,
                                                                 int *__restrict__ err, int cluster)
{
    using TL = Tile<ORDER>;
    constexpr int H = TL::H, GT = TL::GT, JT = TL::JT, S = TL::S;
    constexpr int GT3 = ND == 3 ? GT * GT * GT : GT * GT;
    // 3D field tile x-stride: 68 doubles instead of 8*8, so that the four
    // x-planes of cells a wavefront covers do not share LDS banks
    constexpr int GSX = ND == 3 ? (ORDER == 2 ? 68 : GT * GT) : GT, GF = ND == 3 ? GT * GSX : GT3;
    constexpr int JT3 = ND == 3 ? JT * JT * JT : JT * JT;
    // 3D order 2: J rows padded by one word, so that the tile atomics of a
    // wavefront meet fewer LDS bank conflicts (the strides of deposit3_tile)
    constexpr int JSY = ND == 3 && ORDER == 2 ? JT + 1 : JT, JSX = JT * JSY, JSC = ND == 3 ? JT * JSX : JT3;
    __shared__ double sF[6 * GF];
    __shared__ double sJ[DEPOSIT ? 3 * JSC : 1];
    __shared__ int sHist[kMaxMerged][64];  // destination-cell counts of this cluster's own cells
    // 3D deposit: particles whose node changes on some axis are deposited in
    // a second pass over this queue (old delta, new position in cells, charge
    // weight, old node relative to the cluster), so that the main loop runs
    // the narrow stencil with all lanes.  Order 4: a short queue, so that the
    // workgroup's LDS (field and J tiles of 10^3 and 11^3 points) stays within
    // 80 KB and two workgroups fit a CU.
    constexpr bool QUEUE = DEPOSIT && ND == 3;
    constexpr int QCAP = QUEUE ? (ORDER == 4 ? 12 : 96) : 1;
    __shared__ double sQ[QCAP][8];
    __shared__ int sQn;

    const Cluster cl = decode_cluster(g, cluster);
    int total = 0;
    for (int q = 0; q < ns; ++q)
        total += specs[q].cell_off[cl.gbase + g.cells_per_cl] - specs[q].cell_off[cl.gbase];
    if (total == 0) return;  // uniform across the workgroup

    // ---- stage E, B_m of the cluster + halo
    const double *src[6] = {F.Ex, F.Ey, F.Ez, F.Bx, F.By, F.Bz};
    for (int t = threadIdx.x; t < GT3; t += blockDim.x) {
        int a, b, c;
        if constexpr (ND == 3) {
            a = t / (GT * GT);
            b = (t / GT) % GT;
            c = t % GT;
        } else {
            a = t / GT;
            b = t % GT;
            c = 0;
        }
        const int gi = wrapi(cl.G0[0] - TL::GO + a, g.N[0]);
        const int gj = wrapi(cl.G0[1] - TL::GO + b, g.N[1]);
        const int gk = ND == 3 ? wrapi(cl.G0[2] - TL::GO + c, g.N[2]) : 0;
        const int idx = (gi * g.N[1] + gj) * g.N[2] + gk;  // fewer than 2^27 cells
        const int so = ND == 3 ? a * GSX + b * GT + c : t;
#pragma unroll
        for (int f = 0; f < 6; ++f) sF[f * GF + so] = src[f][idx];
    }
    if constexpr (DEPOSIT)
        for (int t = threadIdx.x; t < 3 * JSC; t += blockDim.x) sJ[t] = 0.;
    for (int t = threadIdx.x; t < kMaxMerged * 64; t += blockDim.x) sHist[t / 64][t % 64] = 0;
    if (threadIdx.x == 0) sQn = 0;
    __syncthreads();

    double pmin[3], pmax[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        pmin[d] = (double)cl.pc[d] * g.plen[d];
        pmax[d] = (double)(cl.pc[d] + 1) * g.plen[d];
    }

    // A wavefront takes 64 consecutive slots of the cluster's range; stored
    // lane-interleaved by the sort, they hold particles of different cells,
    // so the lanes' LDS atomics rarely hit the same address.
    for (int sid = 0; sid < ns; ++sid) {
    const FpSpecies &Sp = specs[sid];
    const SpeciesK &sp = Sp.k;
    const GPart A = glb(Sp.A), B = glb(Sp.B);
    const gptr<const int32_t> perm = glb(Sp.perm);
    const gptr<uint32_t> keys = glb(Sp.keys);
    int32_t *__restrict__ counts = Sp.counts;
    const gptr<int32_t> rank = glb(Sp.rank);
    const bool stay = Sp.stay != nullptr;
    const int lay = Sp.lay;
    const gptr<double> scratch = glb(Sp.scratch);
    const int64_t cap = Sp.cap;
    const int s = Sp.cell_off[cl.gbase], e = Sp.cell_off[cl.gbase + g.cells_per_cl];
    const int N = e - s;
    // consecutive slots per wavefront (lanes in different cells when the
    // sort stored the cluster lane-interleaved)
    for (int rel = threadIdx.x; rel < N; rel += blockDim.x) {
        const int jslot = s + rel;
        const int i = perm[jslot];
        double x[3] = {A.x[i], A.y[i], ND == 3 ? A.z[i] : 0.};
        double u[3] = {A.px[i], A.py[i], A.pz[i]};
        const double w = A.w[i];

        // ---- gather (Interpolator*::coeffs + compute)
        double pn[3], cp[3][5], cd[3][5], dold[3];
        int ip[3], id[3];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            pn[d] = x[d] * g.dinv[d];
            ip[d] = (int)round(pn[d]);
            dold[d] = pn[d] - (double)ip[d];
            spline<ORDER>(dold[d], cp[d]);
            id[d] = (int)round(pn[d] + 0.5);
            spline<ORDER>(pn[d] - (double)id[d] + 0.5, cd[d]);
        }
        int lp[3], ld[3];
        bool inside = true;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            lp[d] = ip[d] - (cl.G0[d] - TL::GO);
            ld[d] = id[d] - (cl.G0[d] - TL::GO);
            inside &= lp[d] >= H && lp[d] <= GT - 1 - H && ld[d] >= H && ld[d] <= GT - 1 - H;
        }
        if (!inside) {
            // not in this cluster (corrupt state): leave it untouched, flag it
            atomicOr(err, 1);
            keys[jslot] = ((stay ? kPlaceAtomic : 13u) << 27) | (uint32_t)(cl.gbase);
            atomicAdd(&counts[cl.gbase], 1);
            B.x[jslot] = x[0];
            B.y[jslot] = x[1];
            if constexpr (ND == 3) B.z[jslot] = x[2];
            B.px[jslot] = u[0];
            B.py[jslot] = u[1];
            B.pz[jslot] = u[2];
            B.w[jslot] = w;
            if constexpr (!DEPOSIT) {
                for (int d = 0; d < 3; ++d) {
                    scratch[d * cap + jslot] = d < ND ? x[d] : 0.;
                    scratch[(3 + d) * cap + jslot] = d < ND ? x[d] : 0.;
                }
                scratch[6 * cap + jslot] = 1.0;
            }
            continue;
        }
        double Ep[3], Bp[3];
        if constexpr (ND == 3) {
            Ep[0] = gather3<H, GT, GSX>(cd[0], cp[1], cp[2], sF + 0 * GF, ld[0], lp[1], lp[2]);
            Ep[1] = gather3<H, GT, GSX>(cp[0], cd[1], cp[2], sF + 1 * GF, lp[0], ld[1], lp[2]);
            Ep[2] = gather3<H, GT, GSX>(cp[0], cp[1], cd[2], sF + 2 * GF, lp[0], lp[1], ld[2]);
            Bp[0] = gather3<H, GT, GSX>(cp[0], cd[1], cd[2], sF + 3 * GF, lp[0], ld[1], ld[2]);
            Bp[1] = gather3<H, GT, GSX>(cd[0], cp[1], cd[2], sF + 4 * GF, ld[0], lp[1], ld[2]);
            Bp[2] = gather3<H, GT, GSX>(cd[0], cd[1], cp[2], sF + 5 * GF, ld[0], ld[1], lp[2]);
        } else {
            Ep[0] = gather2<H, GT>(cd[0], cp[1], sF + 0 * GF, ld[0], lp[1]);
            Ep[1] = gather2<H, GT>(cp[0], cd[1], sF + 1 * GF, lp[0], ld[1]);
            Ep[2] = gather2<H, GT>(cp[0], cp[1], sF + 2 * GF, lp[0], lp[1]);
            Bp[0] = gather2<H, GT>(cp[0], cd[1], sF + 3 * GF, lp[0], ld[1]);
            Bp[1] = gather2<H, GT>(cd[0], cp[1], sF + 4 * GF, ld[0], lp[1]);
            Bp[2] = gather2<H, GT>(cd[0], cd[1], sF + 5 * GF, ld[0], ld[1]);
        }

        // ---- push
        double dxp[3], invgf;
        boris(sp.qmh, g.dt, Ep, Bp, u, dxp, invgf);
        double xn[3];
#pragma unroll
        for (int d = 0; d < ND; ++d) xn[d] = x[d] + dxp[d];

        // ---- boundary tag, periodic wrap, destination key
        int m[3] = {0, 0, 0}, dpc[3] = {cl.pc[0], cl.pc[1], cl.pc[2]};
        double xw[3] = {xn[0], ND > 1 ? xn[1] : 0., ND > 2 ? xn[2] : 0.};
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            if (xn[d] < pmin[d]) m[d] = -1;
            else if (xn[d] >= pmax[d]) m[d] = 1;
            dpc[d] = wrapi(cl.pc[d] + m[d], g.npatch[d]);
            if (m[d] == -1 && cl.pc[d] == 0 && xw[d] < 0.) xw[d] += g.xmax[d];
            if (m[d] == 1 && cl.pc[d] == g.npatch[d] - 1 && xw[d] >= g.xmax[d]) xw[d] -= g.xmax[d];
        }
        const int dp = (dpc[0] * g.npatch[1] + dpc[1]) * g.npatch[2] + dpc[2];
        const uint32_t gkey = (uint32_t)((int64_t)dp * g.cpp + cell_key<ND>(g, dpc, xw));
        // histogram of destination cells for the sort: own-cluster cells in
        // LDS, particles leaving the cluster straight to the global counts.
        // Top key bits: the direction of motion, or with `stay` the rank
        // among this cluster's stayers of the cell (placed by the sort).
        uint32_t tag = (uint32_t)((m[0] + 1) * 9 + (m[1] + 1) * 3 + (m[2] + 1));
        const int64_t rel_cell = (int64_t)gkey - cl.gbase;
        if (rel_cell >= 0 && rel_cell < g.cells_per_cl) {
            if (stay) {
                const int r = atomicAdd(&sHist[sid][rel_cell], 1);
                tag = r < (int)kPlaceMax ? (uint32_t)r : kPlaceAtomic;
            } else {
                atomicAdd(&sHist[sid][rel_cell], 1);
            }
        } else {
            atomicAdd(&counts[gkey], 1);
            if (stay) tag = kPlaceAtomic;
            // cell-major position of this input particle, the tie-break of
            // the canonical in-cell order at the next sort: only read for
            // the particles that leave their cluster (the sort computes it
            // for the others)
            if (rank) rank[jslot] = s + (lay ? cluster_rank(rel, N) : rel);  // layouts 0 and 1 (FP64 form)
        }
        keys[jslot] = (tag << 27) | gkey;
        B.x[jslot] = xw[0];
        B.y[jslot] = xw[1];
        if constexpr (ND == 3) B.z[jslot] = xw[2];
        B.px[jslot] = u[0];
        B.py[jslot] = u[1];
        B.pz[jslot] = u[2];
        B.w[jslot] = w;

        if constexpr (!DEPOSIT) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                scratch[d * cap + jslot] = d < ND ? x[d] : 0.;
                scratch[(3 + d) * cap + jslot] = d < ND ? xn[d] : 0.;
            }
            scratch[6 * cap + jslot] = invgf;
        } else {
            // ---- Esirkepov currents into the LDS tile
            const double cw = sp.cw_scale * w;
            if constexpr (ND == 3) {
                double pnn[3];
                int moved = 0;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    pnn[d] = xn[d] * g.dinv[d];
                    moved |= (int)round(pnn[d]) - ip[d];
                }
                if (moved == 0) deposit3_tile<ORDER, ORDER + 1>(sJ, g, cl.G0, dold, pnn, ip, cw);
                const uint64_t gm = __ballot(moved != 0);
                if (gm) {
                    const int lane = threadIdx.x & 63;
                    const int leader = __ffsll((unsigned long long)gm) - 1;
                    int base = 0;
                    if (lane == leader) base = atomicAdd(&sQn, __popcll(gm));
                    base = __shfl(base, leader, 64);
                    if (moved != 0) {
                        const int pos = base + __popcll(gm & ((1ull << lane) - 1));
                        if (pos < QCAP) {
#pragma unroll
                            for (int d = 0; d < 3; ++d) {
                                sQ[pos][d] = dold[d];
                                sQ[pos][3 + d] = pnn[d];
                            }
                            sQ[pos][6] = cw;
                            sQ[pos][7] = (double)((ip[0] - cl.G0[0]) | ((ip[1] - cl.G0[1]) << 4) |
                                                  ((ip[2] - cl.G0[2]) << 8));
                        } else {
                            deposit3_tile<ORDER, ORDER + 2>(sJ, g, cl.G0, dold, pnn, ip, cw);
                        }
                    }
                }
            } else {
                double S0[3][7], S1[3][7], DS[3][7];
#pragma unroll
                for (int d = 0; d < ND; ++d)
                    esirkepov_shapes<ORDER>(dold[d], xn[d] * g.dinv[d], ip[d], S0[d], S1[d], DS[d]);
                const double crx = cw * (g.dx[0] / g.dt), cry = cw * (g.dx[1] / g.dt);
                const double crz = cw * kThird * u[2] * invgf;
                const int o0 = ip[0] - 2 - (cl.G0[0] - TL::JO);
                const int o1 = ip[1] - 2 - (cl.G0[1] - TL::JO);
                esirkepov2d(S0, S1, DS, crx, cry, crz, [&](int c, int a, int b, double v) {
                    if (v != 0.) atomicAdd(&sJ[c * JT3 + (o0 + a) * JT + (o1 + b)], v);
                });
            }
        }
    }

    }  // species

    __syncthreads();
    for (int t = threadIdx.x; t < ns * 64; t += blockDim.x) {
        const int c = t % 64 < g.cells_per_cl ? sHist[t / 64][t % 64] : 0;
        if (c) {
            const FpSpecies &S = specs[t / 64];
            atomicAdd(&S.counts[cl.gbase + t % 64], c);
            if (S.stay) S.stay[cl.gbase + t % 64] = (uint8_t)(c < (int)kPlaceMax ? c : (int)kPlaceMax);
        }
    }
    if constexpr (QUEUE) {
        const int nq = sQn < QCAP ? sQn : QCAP;
        for (int t = threadIdx.x; t < nq; t += blockDim.x) {
            const double dold[3] = {sQ[t][0], sQ[t][1], sQ[t][2]};
            const double pnn[3] = {sQ[t][3], sQ[t][4], sQ[t][5]};
            const int packed = (int)sQ[t][7];
            const int ip[3] = {cl.G0[0] + (packed & 15), cl.G0[1] + ((packed >> 4) & 15),
                               cl.G0[2] + ((packed >> 8) & 15)};
            deposit3_tile<ORDER, ORDER + 2>(sJ, g, cl.G0, dold, pnn, ip, sQ[t][6]);
        }
        __syncthreads();
    }
    if constexpr (DEPOSIT) {
        // one tile position per thread and pass, its grid index computed once
        // for the three components (32-bit: the engine holds fewer than 2^27
        // cells)
        const gptr<double> dstJ[3] = {glb(J.Jx), glb(J.Jy), glb(J.Jz)};
        for (int r = threadIdx.x; r < JT3; r += blockDim.x) {
            int a, b, k;
            if constexpr (ND == 3) {
                a = r / (JT * JT);
                b = (r / JT) % JT;
                k = r % JT;
            } else {
                a = r / JT;
                b = r % JT;
                k = 0;
            }
            const int gi = wrapi(cl.G0[0] - TL::JO + a, g.N[0]);
            const int gj = wrapi(cl.G0[1] - TL::JO + b, g.N[1]);
            const int gk = ND == 3 ? wrapi(cl.G0[2] - TL::JO + k, g.N[2]) : 0;
            const int idx = (gi * g.N[1] + gj) * g.N[2] + gk;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double v = ND == 3 ? sJ[c * JSC + a * JSX + b * JSY + k] : sJ[c * JT3 + r];
                if (v != 0.) unsafeAtomicAdd((double *)(dstJ[c] + idx), v);
            }
        }
    }
}

// One workgroup per cluster of the grid (xcd: XCD-aware cluster order).
template <int ND, int ORDER, bool DEPOSIT>
__global__ void __launch_bounds__(ND == 3 && ORDER == 4 ? 256 : 512, ND == 3 && ORDER == 4 ? 2 : 3)
k_particles(Geom g, const FpSpecies *__restrict__ specs, int ns, GatherFields F, JPtrs J, int *__restrict__ err,
            int xcd)
{
    fp_cluster<ND, ORDER, DEPOSIT>(g, specs, ns, F, J, err, xcd ? xcd_cluster(blockIdx.x, gridDim.x) : blockIdx.x);
}

// Persistent form: a grid sized to the resident workgroups pulls clusters
// from the work list (the non-empty clusters, wcount[0] of them) through the
// queue head wcount[1]; every workgroup leaves once the list is exhausted.
template <int ND, int ORDER, bool DEPOSIT>
__global__ void __launch_bounds__(ND == 3 && ORDER == 4 ? 256 : 512, ND == 3 && ORDER == 4 ? 2 : 3)
k_particles_list(Geom g, const FpSpecies *__restrict__ specs, int ns, GatherFields F, JPtrs J, int *__restrict__ err,
                 const int *__restrict__ work, int *__restrict__ wcount)
{
    __shared__ int sItem;
    const int nwork = wcount[0];
    for (;;) {
        __syncthreads();  // the previous cluster is done with the LDS tiles
        if (threadIdx.x == 0) sItem = atomicAdd(&wcount[1], 1);
        __syncthreads();
        const int item = sItem;
        if (item >= nwork) break;
        fp_cluster<ND, ORDER, DEPOSIT>(g, specs, ns, F, J, err, work[item]);
    }
}

// Work list of the clusters holding particles of any species of the launch
// (order irrelevant); wcount[0..1] must be zero.
__global__ void k_fp_worklist(Geom g, const FpSpecies *__restrict__ specs, int ns, int *__restrict__ work,
                              int *__restrict__ wcount)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 0;
    if (c < g.npatches * g.clpp) {
        const int64_t gbase = (int64_t)(c / g.clpp) * g.cpp + (int64_t)(c % g.clpp) * g.cells_per_cl;
        for (int q = 0; q < ns; ++q) total += specs[q].cell_off[gbase + g.cells_per_cl] - specs[q].cell_off[gbase];
    }
    const uint64_t m = __ballot(total > 0);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&wcount[0], __popcll(m));
    base = __shfl(base, leader, 64);
    if (total > 0) work[base + __popcll(m & ((1ull << lane) - 1))] = c;
}

// ---------------------------------------------------------------------------
// Deterministic deposit (reference order) into per-patch ghosted J.
// ---------------------------------------------------------------------------
struct PJLayout {
    int dims[3][3];   // dims of Jx, Jy, Jz
    int64_t off[3];   // offsets of Jx, Jy, Jz inside a patch block
    int64_t per;      // doubles per patch
};

__host__ __device__ inline PJLayout pj_layout(const Geom &g)
{
    PJLayout L;
    int np[3];
    for (int d = 0; d < 3; ++d) np[d] = d < g.nd ? g.n[d] + 2 * g.os + 1 : 1;
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) L.dims[c][d] = np[d] + ((d == c && d < g.nd) ? 1 : 0);
    int64_t o = 0;
    for (int c = 0; c < 3; ++c) {
        L.off[c] = o;
        o += (int64_t)L.dims[c][0] * L.dims[c][1] * L.dims[c][2];
    }
    L.per = o;
    return L;
}

// One wavefront per patch.  Each particle's contributions are split over the
// lanes by (component, transverse column); every J point receives at most
// one add per particle, and particles are processed strictly in order: the
// adds are L2 atomics and the barrier after each particle waits for them, so
// every J point sums its contributions in the patch's particle order.
template <int ND, int ORDER>
__global__ void __launch_bounds__(64) k_det_deposit(Geom g, const DetSpecies *__restrict__ species,
                                                    int nspecies, double *__restrict__ pj, int *__restrict__ err)
{
    constexpr int H = ORDER / 2, S = 2 * H + 3;
    const int p = blockIdx.x;
    const PJLayout L = pj_layout(g);
    double *Jc[3] = {pj + p * L.per + L.off[0], pj + p * L.per + L.off[1], pj + p * L.per + L.off[2]};
    int pc[3];
    pc[0] = p / (g.npatch[1] * g.npatch[2]);
    pc[1] = (p / g.npatch[2]) % g.npatch[1];
    pc[2] = p % g.npatch[2];
    int dom[3];
    for (int d = 0; d < 3; ++d) dom[d] = d < ND ? pc[d] * g.n[d] - g.os : 0;
    const int lane = threadIdx.x;

    for (int s = 0; s < nspecies; ++s) {
        const DetSpecies ds = species[s];
        const int a0 = ds.cell_off[(int64_t)p * g.cpp], a1 = ds.cell_off[(int64_t)(p + 1) * g.cpp];
        for (int j = a0; j < a1; ++j) {
            double xo[3], xn[3];
            for (int d = 0; d < 3; ++d) {
                xo[d] = ds.scratch[d * ds.cap + j];
                xn[d] = ds.scratch[(3 + d) * ds.cap + j];
            }
            const double cw = ds.sp.cw_scale * ds.w[j];
            double S0[3][7], S1[3][7], DS[3][7];
            int io[3];
            bool ok = true;
            for (int d = 0; d < ND; ++d) {
                const double pn = xo[d] * g.dinv[d];
                const int ipo = (int)round(pn);
                io[d] = ipo - dom[d];
                const int ipn = (int)round(xn[d] * g.dinv[d]);
                // old node within the patch, move below one cell
                ok &= io[d] >= g.os && io[d] <= g.os + g.n[d] && ipn - ipo >= -1 && ipn - ipo <= 1;
                esirkepov_shapes<ORDER>(pn - (double)ipo, xn[d] * g.dinv[d], ipo, S0[d], S1[d], DS[d]);
            }
            if (!ok) {
                if (lane == 0) atomicOr(err, 1);
                __syncthreads();
                continue;
            }
            if constexpr (ND == 3) {
                const double cr[3] = {cw * (g.dx[0] / g.dt), cw * (g.dx[1] / g.dt), cw * (g.dx[2] / g.dt)};
                // task = component*S*S + column; column (b,c) of the component
                for (int task = lane; task < 3 * S * S; task += 64) {
                    const int c = task / (S * S), col = task % (S * S);
                    const int u = col / S, v = col % S;
                    const int *dm = L.dims[c];
                    double acc = 0.;
                    double W;
                    if (c == 0) W = weight3(S0[1], DS[1], S0[2], DS[2], u, v);
                    else if (c == 1) W = weight3(S0[2], DS[2], S0[0], DS[0], v, u);
                    else W = weight3(S0[0], DS[0], S0[1], DS[1], u, v);
                    const double t = cr[c] * W;
                    for (int r = 1; r < S; ++r) {
                        if constexpr (ORDER == 2) acc -= DS[c][r - 1] * t;
                        else acc -= cr[c] * DS[c][r - 1] * W;
                        int ii, jj, kk;
                        if (c == 0) { ii = r; jj = u; kk = v; }       // Jx(i=r, y=u, z=v)
                        else if (c == 1) { ii = u; jj = r; kk = v; }  // Jy(x=u, j=r, z=v)
                        else { ii = u; jj = v; kk = r; }              // Jz(x=u, y=v, k=r)
                        ii += io[0] - (H + 1);
                        jj += io[1] - (H + 1);
                        kk += io[2] - (H + 1);
                        unsafeAtomicAdd(&Jc[c][((int64_t)ii * dm[1] + jj) * dm[2] + kk], acc);
                    }
                }
            } else {
                const double crx = cw * (g.dx[0] / g.dt), cry = cw * (g.dx[1] / g.dt);
                const double crz = cw * kThird * ds.pz[j] * ds.scratch[6 * ds.cap + j];
                // lanes: 0..4 Jx columns j, 5..9 Jy rows i, 10..34 Jz points
                if (lane < 5) {
                    const int jj = lane;
                    double acc = 0.;
                    for (int i = 1; i < 5; ++i) {
                        if (jj == 0) acc -= crx * DS[0][i - 1] * (0.5 * DS[1][0]);
                        else acc -= crx * DS[0][i - 1] * (S0[1][jj] + 0.5 * DS[1][jj]);
                        const int ii = io[0] - 2 + i, yy = io[1] - 2 + jj;
                        unsafeAtomicAdd(&Jc[0][(int64_t)ii * L.dims[0][1] + yy], acc);
                    }
                } else if (lane < 10) {
                    const int ii = lane - 5;
                    const double ty = S0[0][ii] + 0.5 * DS[0][ii];
                    double acc = 0.;
                    for (int jj = 1; jj < 5; ++jj) {
                        acc -= cry * DS[1][jj - 1] * ty;
                        const int xx = io[0] - 2 + ii, yy = io[1] - 2 + jj;
                        unsafeAtomicAdd(&Jc[1][(int64_t)xx * L.dims[1][1] + yy], acc);
                    }
                } else if (lane < 35) {
                    const int ii = (lane - 10) / 5, jj = (lane - 10) % 5;
                    double h2, h3;
                    if (ii == 0) {
                        h2 = 0.5 * S1[0][0];
                        h3 = S1[0][0];
                    } else {
                        h2 = 0.5 * S1[0][ii] + S0[0][ii];
                        h3 = 0.5 * S0[0][ii] + S1[0][ii];
                    }
                    const double v = jj == 0 ? crz * (S1[1][0] * h3) : crz * (S0[1][jj] * h2 + S1[1][jj] * h3);
                    const int xx = io[0] - 2 + ii, yy = io[1] - 2 + jj;
                    unsafeAtomicAdd(&Jc[2][(int64_t)xx * L.dims[2][1] + yy], v);
                }
            }
            __syncthreads();
        }
    }
}

// Ghost sum along dimension d for the three components:
// low[n + a] += this[a]; this[a] = low[n + a], a in [0, 1 + 2*os + dual).
__global__ void k_ghost_sum(Geom g, int d, double *__restrict__ pj)
{
    const PJLayout L = pj_layout(g);
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // enumerate (patch, component, point in the plane slab)
    int64_t slab[3];
    for (int c = 0; c < 3; ++c) {
        const int gsp = 1 + 2 * g.os + (c == d ? 1 : 0);
        int64_t sz = 1;
        for (int e = 0; e < 3; ++e) sz *= (e == d ? gsp : L.dims[c][e]);
        slab[c] = sz;
    }
    const int64_t per_patch = slab[0] + slab[1] + slab[2];
    if (t >= per_patch * g.npatches) return;
    const int p = (int)(t / per_patch);
    int64_t r = t % per_patch;
    int c = 0;
    while (r >= slab[c]) {
        r -= slab[c];
        ++c;
    }
    const int gsp = 1 + 2 * g.os + (c == d ? 1 : 0);
    int ext[3];
    for (int e = 0; e < 3; ++e) ext[e] = e == d ? gsp : L.dims[c][e];
    int idx[3];
    idx[2] = (int)(r % ext[2]);
    idx[1] = (int)((r / ext[2]) % ext[1]);
    idx[0] = (int)(r / ((int64_t)ext[2] * ext[1]));
    int pc[3];
    pc[0] = p / (g.npatch[1] * g.npatch[2]);
    pc[1] = (p / g.npatch[2]) % g.npatch[1];
    pc[2] = p % g.npatch[2];
    pc[d] = wrapi(pc[d] - 1, g.npatch[d]);
    const int lp = (pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2];
    const int *dm = L.dims[c];
    double *me = pj + (int64_t)p * L.per + L.off[c];
    double *lo = pj + (int64_t)lp * L.per + L.off[c];
    int li[3] = {idx[0], idx[1], idx[2]};
    li[d] += g.n[d];
    const int64_t a = ((int64_t)li[0] * dm[1] + li[1]) * dm[2] + li[2];
    const int64_t b = ((int64_t)idx[0] * dm[1] + idx[1]) * dm[2] + idx[2];
    lo[a] += me[b];
    me[b] = lo[a];
}

// Global J[g] = owning patch's ghost-summed value.
__global__ void k_fold_interior(Geom g, const double *__restrict__ pj, JPtrs J)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const PJLayout L = pj_layout(g);
    const int gk = (int)(t % g.N[2]);
    const int gj = (int)((t / g.N[2]) % g.N[1]);
    const int gi = (int)(t / ((int64_t)g.N[2] * g.N[1]));
    const int gg[3] = {gi, gj, gk};
    int pc[3], l[3];
    for (int d = 0; d < 3; ++d) {
        pc[d] = gg[d] / g.n[d];
        l[d] = gg[d] - pc[d] * g.n[d] + (d < g.nd ? g.os : 0);
    }
    const int p = (pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2];
    double *dst[3] = {J.Jx, J.Jy, J.Jz};
    for (int c = 0; c < 3; ++c) {
        const int *dm = L.dims[c];
        dst[c][t] = pj[(int64_t)p * L.per + L.off[c] + ((int64_t)l[0] * dm[1] + l[1]) * dm[2] + l[2]];
    }
}

}  // namespace

int64_t patch_j_size(const Geom &g) { return pj_layout(g).per; }

void launch_particles(const Geom &g, const FpSpecies *specs, int ns, const GatherFields &F, const JPtrs &J,
                      bool deposit, int *err, int threads, bool xcd, int persistent_blocks, int *work, int *wcount,
                      hipStream_t st)
{
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
        throw Error(-1, "particle-kernel workgroup size must be 64, 128, 256 or 512");
    if (threads > particle_max_threads(g, false))
        throw Error(-1, "workgroup size above the particle kernel's limit for this geometry and order");
    if (ns < 1 || ns > kMaxMerged) throw Error(-1, "bad species count of a merged particle launch");
    const int nb = g.npatches * g.clpp;
    const dim3 block(threads);
    if (persistent_blocks > 0) {
        // the work list, then as many workgroups as fit at once
        launch_fill_i32(wcount, 2, 0, st);
        hipLaunchKernelGGL(k_fp_worklist, dim3((nb + 255) / 256), dim3(256), 0, st, g, specs, ns, work, wcount);
        SMI_LAUNCH_CHECK();
    }
    const dim3 grid(persistent_blocks > 0 ? std::min(nb, persistent_blocks) : nb);
#define SMI_PK(ND, O, D)                                                                                          \
    do {                                                                                                          \
        if (persistent_blocks > 0)                                                                                \
            hipLaunchKernelGGL((k_particles_list<ND, O, D>), grid, block, 0, st, g, specs, ns, F, J, err,         \
                               (const int *)work, wcount);                                                         \
        else                                                                                                      \
            hipLaunchKernelGGL((k_particles<ND, O, D>), grid, block, 0, st, g, specs, ns, F, J, err, xcd ? 1 : 0); \
    } while (0)
    if (g.nd == 3 && g.order == 2) {
        if (deposit) SMI_PK(3, 2, true);
        else SMI_PK(3, 2, false);
    } else if (g.nd == 3 && g.order == 4) {
        if (deposit) SMI_PK(3, 4, true);
        else SMI_PK(3, 4, false);
    } else if (g.nd == 2 && g.order == 2) {
        if (deposit) SMI_PK(2, 2, true);
        else SMI_PK(2, 2, false);
    } else {
        throw Error(-1, "unsupported (ndim, order) for the particle kernel");
    }
#undef SMI_PK
    SMI_LAUNCH_CHECK();
}

int fp_persistent_blocks(const Geom &g, bool deposit, int threads)
{
    int dev = 0, cus = 0, per_cu = 0;
    SMI_HIP(hipGetDevice(&dev));
    SMI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void *fn = nullptr;
    if (g.nd == 3 && g.order == 2)
        fn = deposit ? (const void *)k_particles_list<3, 2, true> : (const void *)k_particles_list<3, 2, false>;
    else if (g.nd == 3 && g.order == 4)
        fn = deposit ? (const void *)k_particles_list<3, 4, true> : (const void *)k_particles_list<3, 4, false>;
    else
        fn = deposit ? (const void *)k_particles_list<2, 2, true> : (const void *)k_particles_list<2, 2, false>;
    SMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
    return std::max(1, cus * std::max(per_cu, 1));
}

int particle_max_threads(const Geom &g, bool mixed)
{
    // the 3D order-4 FP64 kernel needs the full register file of one wave
    return !mixed && g.nd == 3 && g.order == 4 ? 256 : 512;
}

void launch_det_deposit(const Geom &g, const DetSpecies *species, int nspecies, double *pj, int *err,
                        hipStream_t st)
{
    const PJLayout L = pj_layout(g);
    launch_fill(pj, L.per * g.npatches, 0.0, st);
    if (g.nd == 3 && g.order == 2)
        hipLaunchKernelGGL((k_det_deposit<3, 2>), dim3(g.npatches), dim3(64), 0, st, g, species, nspecies, pj, err);
    else if (g.nd == 3 && g.order == 4)
        hipLaunchKernelGGL((k_det_deposit<3, 4>), dim3(g.npatches), dim3(64), 0, st, g, species, nspecies, pj, err);
    else if (g.nd == 2 && g.order == 2)
        hipLaunchKernelGGL((k_det_deposit<2, 2>), dim3(g.npatches), dim3(64), 0, st, g, species, nspecies, pj, err);
    else
        throw Error(-1, "unsupported (ndim, order) for the deterministic deposit");
    SMI_LAUNCH_CHECK();
}

void launch_ghost_sum(const Geom &g, double *pj, hipStream_t st)
{
    const PJLayout L = pj_layout(g);
    for (int d = 0; d < g.nd; ++d) {
        int64_t per = 0;
        for (int c = 0; c < 3; ++c) {
            const int gsp = 1 + 2 * g.os + (c == d ? 1 : 0);
            int64_t sz = 1;
            for (int e = 0; e < 3; ++e) sz *= (e == d ? gsp : L.dims[c][e]);
            per += sz;
        }
        const int64_t tot = per * g.npatches;
        hipLaunchKernelGGL(k_ghost_sum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, g, d, pj);
        SMI_LAUNCH_CHECK();
    }
}

void launch_ghost_sum_and_fold(const Geom &g, double *pj, const JPtrs &J, hipStream_t st)
{
    launch_ghost_sum(g, pj, st);
    hipLaunchKernelGGL(k_fold_interior, dim3((unsigned)((g.ncells + 255) / 256)), dim3(256), 0, st, g,
                       (const double *)pj, J);
    SMI_LAUNCH_CHECK();
}

}  // namespace smi
