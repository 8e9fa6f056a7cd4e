// seams.hip -- per-operator entry points for the remaining reference seams:
//
//  * smi_current_deposition_3d / smi_density_deposition_3d keep the argument
//    lists of the reference's own C seams currentDeposition3DOnDevice and
//    densityDeposition3DOnDevice (src/Projector/Projector3D2OrderGPU.cpp:
//    55-120): one patch, particles binned by 4^3 cluster (bin_index = the
//    clusters' cumulative ends, Particles::last_index), per-particle int16
//    charge, SmileiMPI scratch iold/deltaold with stride n;
//  * smi_density_3d / smi_density_2d: Projector3D2Order / 3D4Order /
//    2D2Order ::basic, type 0 (the charge density of diagnostic steps and of
//    frozen species);
//  * smi_maxwell_ampere_2d / smi_maxwell_faraday_2d (MA_Solver2D_norm,
//    MF_Solver2D_Yee) on one patch;
//  * smi_sum_j_ghosts / smi_exchange_b: SyncVectorPatch::sumAllComponents and
//    exchangeB over all patches of a device, held in one patch-major array;
//  * smi_sort_exchange: the particle exchange between patches and the cell
//    sort of one species (Patch::exchParticles / importAndSortParticles,
//    Species::sortParticles) on one device.
#include <cstdio>
#include <vector>

#include "../../include/smilei_amd.h"
#include "kernels.h"
#include "pic_device.h"
#include "seams_internal.h"

using namespace smi;
using namespace smi::dev;
using namespace smi::seams;

int smi::seams::op_fail(int code, const std::string &m)
{
    set_error(m);
    return code;
}

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

#define SEAM_CHECK()                                                              \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess) return op_fail(SMI_ERR_HIP, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Cluster-tiled deposits of the reference seams.  One workgroup per 4^3
// cluster of the patch (bins in the reference's x-major order); the cluster's
// tile spans the 9 array points [4c, 4c+9) per axis: the order-2 stencils
// (old nearest node +-2) of every particle binned in the cluster.  A particle
// whose stencil would leave the tile (not binned where bin_index says, or
// moved more than one cell) is skipped and flagged.
// ---------------------------------------------------------------------------
constexpr int kCT = 9, kCT3 = kCT * kCT * kCT;

struct SeamArgs {
    double *out[3];     // Jx, Jy, Jz (currents) or rho (density)
    int64_t size[3];    // their sizes in doubles
    const double *pos[3];
    const short *charge;
    const double *weight;
    const int *bin_index;
    int nb[3];
    const int *iold;
    const double *dold;
    int64_t n;          // stride of iold / deltaold
    double icv, dinv[3], dxdt[3];
    int dom[3], nprimy, nprimz, not_spectral, cell_sorting;
    int *err;
};

__device__ inline int bin_coords(const SeamArgs &a, int b, int *bc)
{
    bc[0] = b / (a.nb[1] * a.nb[2]);
    bc[1] = (b / a.nb[2]) % a.nb[1];
    bc[2] = b % a.nb[2];
    return b;
}

// Particle range of bin b, and this thread's share of it: a contiguous chunk
// per thread with cell_sorting (neighbours in a cell land on one thread), a
// strided walk otherwise (Projector3D2OrderGPUKernelCUDAHIP.cu:237-256 picks
// the same two orders).  Only the order of the tile atomics differs.
__device__ inline void bin_share(const SeamArgs &a, int b, int64_t &beg, int64_t &end, int64_t &stride)
{
    int64_t first = b == 0 ? 0 : a.bin_index[b - 1], last = a.bin_index[b];
    first = first < 0 ? 0 : (first > a.n ? a.n : first);
    last = last < first ? first : (last > a.n ? a.n : last);
    if (a.cell_sorting) {
        const int64_t per = (last - first + blockDim.x - 1) / blockDim.x;
        beg = first + (int64_t)threadIdx.x * per;
        end = beg + per < last ? beg + per : last;
        stride = 1;
    } else {
        beg = first + threadIdx.x;
        end = last;
        stride = blockDim.x;
    }
}

__global__ void __launch_bounds__(256) k_seam_currents(SeamArgs a)
{
    __shared__ double sJ[3 * kCT3];
    const int b = blockIdx.x;
    int bc[3];
    bin_coords(a, b, bc);
    for (int t = threadIdx.x; t < 3 * kCT3; t += blockDim.x) sJ[t] = 0.;
    __syncthreads();
    int64_t beg, end, stride;
    bin_share(a, b, beg, end, stride);
    for (int64_t i = beg; i < end; i += stride) {
        double S0[3][7], S1[3][7], DS[3][7];
        int o[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            const int io = a.iold[d * a.n + i];
            const double pn = a.pos[d][i] * a.dinv[d];
            const int sh = (int)round(pn) - (io + a.dom[d]);
            o[d] = io - 2 - 4 * bc[d];  // tile index of stencil point 0
            ok &= sh >= -1 && sh <= 1 && o[d] >= 0 && o[d] <= kCT - 5;
            if (ok) esirkepov_shapes<2>(a.dold[d * a.n + i], pn, io + a.dom[d], S0[d], S1[d], DS[d]);
        }
        if (!ok) {
            atomicOr(a.err, 1);
            continue;
        }
        // charge_weight and cr_p as Projector3D2OrderGPUKernelCUDAHIP.cu:281-284
        const double cw = a.icv * (double)a.charge[i] * a.weight[i];
        const double cr[3] = {cw * a.dxdt[0], cw * a.dxdt[1], cw * a.dxdt[2]};
        esirkepov3d<2>(S0, DS, cr, [&](int c, int u, int v, int k, double val) {
            if (val == 0.) return;
            unsafeAtomicAdd(&sJ[c * kCT3 + ((o[0] + u) * kCT + (o[1] + v)) * kCT + (o[2] + k)], val);
        });
    }
    __syncthreads();
    // flush: Jy and Jz are dual along y / z when not_spectral (their rows are
    // one point longer), as the reference's flush indexing encodes
    for (int t = threadIdx.x; t < 3 * kCT3; t += blockDim.x) {
        const double v = sJ[t];
        if (v == 0.) continue;
        const int c = t / kCT3, r = t % kCT3;
        const int gx = 4 * bc[0] + r / (kCT * kCT), gy = 4 * bc[1] + (r / kCT) % kCT, gz = 4 * bc[2] + r % kCT;
        int64_t idx = ((int64_t)gx * a.nprimy + gy) * a.nprimz + gz;
        if (c == 1) idx += (int64_t)a.not_spectral * gx * a.nprimz;
        if (c == 2) idx += (int64_t)a.not_spectral * ((int64_t)gx * a.nprimy + gy);
        if (idx < 0 || idx >= a.size[c]) {
            atomicOr(a.err, 2);
            continue;
        }
        unsafeAtomicAdd(&a.out[c][idx], v);
    }
}

__global__ void __launch_bounds__(256) k_seam_density(SeamArgs a)
{
    __shared__ double sR[kCT3];
    const int b = blockIdx.x;
    int bc[3];
    bin_coords(a, b, bc);
    for (int t = threadIdx.x; t < kCT3; t += blockDim.x) sR[t] = 0.;
    __syncthreads();
    int64_t beg, end, stride;
    bin_share(a, b, beg, end, stride);
    for (int64_t i = beg; i < end; i += stride) {
        double S0[3][7], S1[3][7], DS[3][7];
        int o[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            const int io = a.iold[d * a.n + i];
            const double pn = a.pos[d][i] * a.dinv[d];
            const int sh = (int)round(pn) - (io + a.dom[d]);
            o[d] = io - 2 - 4 * bc[d];
            ok &= sh >= -1 && sh <= 1 && o[d] >= 0 && o[d] <= kCT - 5;
            // S1: the new position's weights on the 5-point window of the
            // old node (init_S1, Projector3D2OrderGPUKernelCUDAHIP.cu:116-131)
            if (ok) esirkepov_shapes<2>(0., pn, io + a.dom[d], S0[d], S1[d], DS[d]);
        }
        if (!ok) {
            atomicOr(a.err, 1);
            continue;
        }
        const double cw = a.icv * (double)a.charge[i] * a.weight[i];
        for (int u = 0; u < 5; ++u)
            for (int v = 0; v < 5; ++v) {
                const double t = cw * S1[0][u] * S1[1][v];
                for (int k = 0; k < 5; ++k) {
                    const double val = t * S1[2][k];
                    if (val == 0.) continue;
                    unsafeAtomicAdd(&sR[((o[0] + u) * kCT + (o[1] + v)) * kCT + (o[2] + k)], val);
                }
            }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kCT3; t += blockDim.x) {
        const double v = sR[t];
        if (v == 0.) continue;
        const int gx = 4 * bc[0] + t / (kCT * kCT), gy = 4 * bc[1] + (t / kCT) % kCT, gz = 4 * bc[2] + t % kCT;
        const int64_t idx = ((int64_t)gx * a.nprimy + gy) * a.nprimz + gz;
        if (idx < 0 || idx >= a.size[0]) {
            atomicOr(a.err, 2);
            continue;
        }
        unsafeAtomicAdd(&a.out[0][idx], v);
    }
}

// Checks shared by both seams; fills SeamArgs.  Returns SMI_OK or an error.
int seam_args(SeamArgs &a, const int *bin_index, unsigned nbx, unsigned nby, unsigned nbz, const int *iold,
              const double *deltaold, unsigned n, double inv_cell_volume, double dx_inv, double dy_inv,
              double dz_inv, double dx_ov_dt, double dy_ov_dt, double dz_ov_dt, int i_dom, int j_dom, int k_dom,
              int nprimy, int nprimz, int not_spectral, bool cell_sorting)
{
    if (nbx == 0 || nby == 0 || nbz == 0 || (uint64_t)nbx * nby * nbz > (1u << 30))
        return op_fail(SMI_ERR_INVALID, "bin counts must be positive");
    if (nprimy < 9 || nprimz < 9) return op_fail(SMI_ERR_INVALID, "nprimy and nprimz must hold a 9-point tile");
    if (4 * (int64_t)(nby - 1) + kCT > nprimy || 4 * (int64_t)(nbz - 1) + kCT > nprimz)
        return op_fail(SMI_ERR_INVALID, "bin counts do not fit the patch arrays (4 cells per bin + 5 ghost points)");
    if (not_spectral != 0 && not_spectral != 1) return op_fail(SMI_ERR_INVALID, "not_spectral must be 0 or 1");
    if (n > 0 && (!bin_index || !iold || !deltaold)) return op_fail(SMI_ERR_INVALID, "null particle scratch");
    a.bin_index = bin_index;
    a.nb[0] = (int)nbx;
    a.nb[1] = (int)nby;
    a.nb[2] = (int)nbz;
    a.iold = iold;
    a.dold = deltaold;
    a.n = n;
    a.icv = inv_cell_volume;
    a.dinv[0] = dx_inv;
    a.dinv[1] = dy_inv;
    a.dinv[2] = dz_inv;
    a.dxdt[0] = dx_ov_dt;
    a.dxdt[1] = dy_ov_dt;
    a.dxdt[2] = dz_ov_dt;
    a.dom[0] = i_dom;
    a.dom[1] = j_dom;
    a.dom[2] = k_dom;
    a.nprimy = nprimy;
    a.nprimz = nprimz;
    a.not_spectral = not_spectral;
    a.cell_sorting = cell_sorting ? 1 : 0;
    return SMI_OK;
}

// Runs one seam kernel on the null stream with a fresh error flag, waits for
// it (the reference's launcher synchronises the device after the kernel) and
// reports skipped particles.
template <class Kernel>
int run_seam(Kernel k, SeamArgs &a)
{
    int *err = nullptr;
    if (hipMalloc(&err, sizeof(int)) != hipSuccess) return op_fail(SMI_ERR_HIP, "hipMalloc of the error flag failed");
    try {
        launch_fill_i32(err, 1, 0, nullptr);
    } catch (const std::exception &ex) {
        (void)hipFree(err);
        return op_fail(SMI_ERR_HIP, ex.what());
    }
    a.err = err;
    hipLaunchKernelGGL(k, dim3((unsigned)(a.nb[0] * a.nb[1] * a.nb[2])), dim3(256), 0, nullptr, a);
    hipError_t e = hipGetLastError();
    int h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, err, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(err);
    if (e != hipSuccess) return op_fail(SMI_ERR_HIP, hipGetErrorString(e));
    if (h & 1)
        return op_fail(SMI_ERR_INVALID, "a particle's stencil leaves its bin's tile (not binned by bin_index, or moved "
                                        "more than one cell); it was skipped");
    if (h & 2) return op_fail(SMI_ERR_INVALID, "a tile point falls outside the output array (sizes too small)");
    return SMI_OK;
}

// ---------------------------------------------------------------------------
// Projector::basic type 0 (charge density at the particles' positions), one
// thread per particle, FP64 atomics into the patch array.
// ---------------------------------------------------------------------------
struct BasicArgs {
    double *rho;
    int np[3];
    int dom[3];
    double dinv[3];
    double cw_scale;  // inv_cell_volume * charge
    int64_t n;
    const double *pos[3];
    const double *w;
    int *err;
};

template <int ND, int ORDER>
__global__ void __launch_bounds__(256) k_basic(BasicArgs a)
{
    constexpr int H = ORDER / 2, W = 2 * H + 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double S[3][5];
    int o[3] = {0, 0, 0};
    bool ok = true;
    for (int d = 0; d < ND; ++d) {
        const double pn = a.pos[d][i] * a.dinv[d];
        const int ip = (int)round(pn);
        spline<ORDER>(pn - (double)ip, S[d]);
        o[d] = ip - a.dom[d] - H;
        ok &= o[d] >= 0 && o[d] + W <= a.np[d];
    }
    if (!ok) {
        atomicOr(a.err, 1);
        return;
    }
    const double cw = a.cw_scale * a.w[i];
    if constexpr (ND == 3) {
        for (int u = 0; u < W; ++u)
            for (int v = 0; v < W; ++v)
                for (int k = 0; k < W; ++k)
                    unsafeAtomicAdd(&a.rho[((int64_t)(o[0] + u) * a.np[1] + (o[1] + v)) * a.np[2] + (o[2] + k)],
                                    cw * S[0][u] * S[1][v] * S[2][k]);
    } else {
        for (int u = 0; u < W; ++u)
            for (int v = 0; v < W; ++v)
                unsafeAtomicAdd(&a.rho[(int64_t)(o[0] + u) * a.np[1] + (o[1] + v)], cw * S[0][u] * S[1][v]);
    }
}

int run_basic(int nd, int order, BasicArgs &a, hipStream_t st)
{
    int *err = nullptr;
    if (hipMalloc(&err, sizeof(int)) != hipSuccess) return op_fail(SMI_ERR_HIP, "hipMalloc of the error flag failed");
    try {
        launch_fill_i32(err, 1, 0, st);
    } catch (const std::exception &ex) {
        (void)hipFree(err);
        return op_fail(SMI_ERR_HIP, ex.what());
    }
    a.err = err;
    const dim3 grid(nblk(a.n, 256)), block(256);
    if (nd == 3 && order == 2) hipLaunchKernelGGL((k_basic<3, 2>), grid, block, 0, st, a);
    else if (nd == 3) hipLaunchKernelGGL((k_basic<3, 4>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_basic<2, 2>), grid, block, 0, st, a);
    hipError_t e = hipGetLastError();
    int h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(err);
    if (e != hipSuccess) return op_fail(SMI_ERR_HIP, hipGetErrorString(e));
    if (h) return op_fail(SMI_ERR_INVALID, "a particle's stencil leaves the patch arrays; it was skipped");
    return SMI_OK;
}

// ---------------------------------------------------------------------------
// MA_Solver2D_norm / MF_Solver2D_Yee on one patch (same arithmetic and
// association as MA_Solver2D_norm.cpp:54-105, MF_Solver2D_Yee.cpp:50-99).
// ---------------------------------------------------------------------------
__global__ void k_ma2d_patch(double *Ex, double *Ey, double *Ez, const double *Bx, const double *By,
                             const double *Bz, const double *Jx, const double *Jy, const double *Jz, int nxp, int nyp,
                             double dt, double dtx, double dty)
{
    const int64_t nxd = nxp + 1, nyd = nyp + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nEx = nxd * nyp, nEy = (int64_t)nxp * nyd, nEz = (int64_t)nxp * nyp;
    if (t < nEx) {
        const int64_t x = t / nyp, y = t % nyp;
        Ex[t] += -dt * Jx[t] + dty * (Bz[x * nyd + y + 1] - Bz[x * nyd + y]);
    } else if (t < nEx + nEy) {
        const int64_t e = t - nEx, x = e / nyd, y = e % nyd;
        Ey[e] += -dt * Jy[e] - dtx * (Bz[(x + 1) * nyd + y] - Bz[x * nyd + y]);
    } else if (t < nEx + nEy + nEz) {
        const int64_t e = t - nEx - nEy, x = e / nyp, y = e % nyp;
        Ez[e] += -dt * Jz[e] + dtx * (By[(x + 1) * nyp + y] - By[x * nyp + y]) -
                 dty * (Bx[x * nyd + y + 1] - Bx[x * nyd + y]);
    }
}

__global__ void k_mf2d_patch(double *Bx, double *By, double *Bz, const double *Ex, const double *Ey,
                             const double *Ez, int nxp, int nyp, double dtx, double dty)
{
    const int64_t nxd = nxp + 1, nyd = nyp + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nBx = (int64_t)nxp * nyd, nBy = nxd * nyp, nBz = nxd * nyd;
    if (t < nBx) {
        const int64_t x = t / nyd, y = t % nyd;
        if (y < 1 || y >= nyd - 1) return;
        Bx[t] -= dty * (Ez[x * nyp + y] - Ez[x * nyp + y - 1]);
    } else if (t < nBx + nBy) {
        const int64_t e = t - nBx, x = e / nyp, y = e % nyp;
        if (x < 1 || x >= nxd - 1) return;
        By[e] += dtx * (Ez[x * nyp + y] - Ez[(x - 1) * nyp + y]);
    } else if (t < nBx + nBy + nBz) {
        const int64_t e = t - nBx - nBy, x = e / nyd, y = e % nyd;
        if (x < 1 || x >= nxd - 1 || y < 1 || y >= nyd - 1) return;
        Bz[e] += dty * (Ex[x * nyp + y] - Ex[x * nyp + y - 1]) - dtx * (Ey[x * nyd + y] - Ey[(x - 1) * nyd + y]);
    }
}

// ---------------------------------------------------------------------------
// B exchange between the patches of a patch-major array (exchangeB,
// SyncVectorPatch.cpp:667, 1441-1830), one pass per axis d: for each B
// component dual along d, this patch's planes [0, os) <- the low neighbour's
// [n, n+os), and the low neighbour's [n+os+2, n+2os+2) <- this patch's
// [os+2, 2os+2).  Within a pass no point is both read and written.
// ---------------------------------------------------------------------------
struct PatchLayout {
    int dims[3][3];
    int64_t off[3], per;
};

PatchLayout patch_layout(const Geom &g, const int dual[3][3])
{
    PatchLayout L;
    int64_t o = 0;
    for (int c = 0; c < 3; ++c) {
        for (int d = 0; d < 3; ++d)
            L.dims[c][d] = d < g.nd ? g.n[d] + 2 * g.os + 1 + dual[c][d] : 1;
        L.off[c] = o;
        o += (int64_t)L.dims[c][0] * L.dims[c][1] * L.dims[c][2];
    }
    L.per = o;
    return L;
}

const int kDualJ[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
const int kDualB[3][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

__device__ inline int low_patch(const Geom &g, int p, int d)
{
    int pc[3] = {p / (g.npatch[1] * g.npatch[2]), (p / g.npatch[2]) % g.npatch[1], p % g.npatch[2]};
    pc[d] = wrapi(pc[d] - 1, g.npatch[d]);
    return (pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2];
}

__global__ void k_exchange_b(Geom g, PatchLayout L, int d, double *__restrict__ B)
{
    // work items: (patch, component dual along d, point of an os-plane slab)
    int64_t slab[3];
    for (int c = 0; c < 3; ++c) {
        slab[c] = 0;
        if (c == d) continue;  // B_c is dual along every axis but c
        int64_t sz = 1;
        for (int e = 0; e < 3; ++e) sz *= e == d ? g.os : L.dims[c][e];
        slab[c] = sz;
    }
    const int64_t per_patch = slab[0] + slab[1] + slab[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_patch * g.npatches) return;
    const int p = (int)(t / per_patch);
    int64_t r = t % per_patch;
    int c = 0;
    while (r >= slab[c]) {
        r -= slab[c];
        ++c;
    }
    const int *dm = L.dims[c];
    int ext[3];
    for (int e = 0; e < 3; ++e) ext[e] = e == d ? g.os : dm[e];
    int idx[3];
    idx[2] = (int)(r % ext[2]);
    idx[1] = (int)((r / ext[2]) % ext[1]);
    idx[0] = (int)(r / ((int64_t)ext[2] * ext[1]));
    double *me = B + (int64_t)p * L.per + L.off[c];
    double *lo = B + (int64_t)low_patch(g, p, d) * L.per + L.off[c];
    auto at = [&](const int *q) { return ((int64_t)q[0] * dm[1] + q[1]) * dm[2] + q[2]; };
    int a[3] = {idx[0], idx[1], idx[2]};
    a[d] += g.n[d];
    me[at(idx)] = lo[at(a)];
    const int gsp = g.os + 2;
    int e3[3] = {idx[0], idx[1], idx[2]}, f3[3] = {idx[0], idx[1], idx[2]};
    e3[d] += g.n[d] + gsp;
    f3[d] += gsp;
    lo[at(e3)] = me[at(f3)];
}

}  // namespace

// ---------------------------------------------------------------------------
// Geometry of the per-device entry points (the checks of smi_create).
// ---------------------------------------------------------------------------
int smi::seams::make_geom(int ndim, int order, const int n_cells[3], const int patch_size[3], const double *cell_length,
              Geom &g)
{
    if (ndim != 2 && ndim != 3) return op_fail(SMI_ERR_INVALID, "ndim must be 2 or 3");
    if (order != 2 && order != 4) return op_fail(SMI_ERR_INVALID, "order must be 2 or 4");
    if (ndim == 2 && order != 2) return op_fail(SMI_ERR_INVALID, "2D order must be 2");
    g = Geom{};
    g.nd = ndim;
    g.order = order;
    g.os = order;
    int64_t ncells = 1;
    for (int d = 0; d < 3; ++d) {
        if (d < ndim) {
            if (n_cells[d] <= 0 || patch_size[d] <= 0 || !(cell_length[d] > 0))
                return op_fail(SMI_ERR_INVALID, "n_cells, patch_size and cell_length must be positive");
            if (patch_size[d] % 4) return op_fail(SMI_ERR_INVALID, "patch_size must be a multiple of 4");
            if (n_cells[d] % patch_size[d]) return op_fail(SMI_ERR_INVALID, "patch_size must divide n_cells");
            if (patch_size[d] < 2 * g.os + 2)
                return op_fail(SMI_ERR_INVALID, "patch_size must be >= 2*oversize+2 for the ghost sum");
            g.N[d] = n_cells[d];
            g.n[d] = patch_size[d];
            g.dx[d] = cell_length[d];
        } else {
            g.N[d] = 1;
            g.n[d] = 1;
            g.dx[d] = 1.0;
        }
        g.npatch[d] = g.N[d] / g.n[d];
        g.ncl[d] = d < ndim ? g.n[d] / 4 : 1;
        g.dinv[d] = 1.0 / g.dx[d];
        g.plen[d] = (double)g.n[d] * g.dx[d];
        g.xmax[d] = g.dx[d] * (double)g.N[d];
        ncells *= g.N[d];
    }
    if (ncells >= (1ll << 27)) return op_fail(SMI_ERR_INVALID, "at most 2^27 cells");
    g.ncells = ncells;
    g.cpp = g.n[0] * g.n[1] * g.n[2];
    g.clpp = g.ncl[0] * g.ncl[1] * g.ncl[2];
    g.cells_per_cl = ndim == 3 ? 64 : 16;
    g.npatches = g.npatch[0] * g.npatch[1] * g.npatch[2];
    g.dt = 1.0;
    return SMI_OK;
}

// ---------------------------------------------------------------------------
// Particle exchange + cell sort of one species over all patches of a device.
// ---------------------------------------------------------------------------
namespace {

// Periodic wrap of Patch::prepareParticles (Patch.cpp:624-668): a position
// that left the box by less than its length re-enters on the other side;
// one farther out is flagged and left alone.
template <int ND>
__global__ void k_wrap(Geom g, double *x, double *y, double *z, int64_t n, int *err)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *p[3] = {x, y, z};
    for (int d = 0; d < ND; ++d) {
        double v = p[d][i];
        if (v < 0.) v += g.xmax[d];
        else if (v >= g.xmax[d]) v -= g.xmax[d];
        // a tiny negative position can round to xmax itself: kept, and
        // clamped into the last cell by the key like the reference's
        if (!(v >= 0. && v <= g.xmax[d])) {
            atomicOr(err, 1);
            return;
        }
        p[d][i] = v;
    }
}

template <class T>
__global__ void k_gather_perm(const int32_t *__restrict__ perm, const T *__restrict__ src, T *__restrict__ dst,
                              int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// In-patch cell key of sorted slot i.
__global__ void k_keys_out(Geom g, const int32_t *__restrict__ perm, const uint32_t *__restrict__ keys,
                           int32_t *__restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)((keys[perm[i]] & ((1u << 27) - 1)) % (uint32_t)g.cpp);
}

// Per-patch particle counts and, per patch, the cumulative ends of its
// clusters relative to the patch's first particle (Particles::last_index).
__global__ void k_bins_out(Geom g, const int32_t *__restrict__ cell_off, int32_t *__restrict__ patch_count,
                           int32_t *__restrict__ bin_index)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)g.npatches * g.clpp) return;
    const int p = (int)(t / g.clpp), b = (int)(t % g.clpp);
    const int64_t c0 = (int64_t)p * g.cpp;
    if (bin_index) bin_index[t] = cell_off[c0 + (int64_t)(b + 1) * g.cells_per_cl] - cell_off[c0];
    if (patch_count && b == 0) patch_count[p] = cell_off[c0 + g.cpp] - cell_off[c0];
}

constexpr int64_t kAlignWords = 64;

int64_t align_words(int64_t w) { return (w + kAlignWords - 1) / kAlignWords * kAlignWords; }

}  // namespace

int64_t smi::seams::sort_scratch_words(const Geom &g, int64_t n)
{
    const int64_t nc = g.ncells, m = n > 0 ? n : 1;
    return align_words(nc) + align_words(nc + 1) + 2 * align_words(m) + align_words(m / 17 + 1) + align_words(2) +
           align_words(sort_bsum_size(nc)) + align_words((nc + 3) / 4) + align_words(nc / g.cells_per_cl) +
           align_words(m) + align_words(2 * m) + align_words(1);
}

SortScratch smi::seams::carve_sort_scratch(const Geom &g, int64_t n, int32_t *base)
{
    const int64_t nc = g.ncells, m = n > 0 ? n : 1;
    SortScratch s{};
    int32_t *p = base;
    auto take = [&](int64_t words) {
        int32_t *r = p;
        p += align_words(words);
        return r;
    };
    s.sb.counts = take(nc);
    s.sb.cell_off = take(nc + 1);
    s.sb.perm = take(m);
    s.sb.tmp = take(m);
    s.sb.big = take(m / 17 + 1);
    s.sb.nbig = take(2);
    s.sb.nbsum = sort_bsum_size(nc);
    s.sb.bsum = take(s.sb.nbsum);
    s.sb.stay = (uint8_t *)take((nc + 3) / 4);
    s.sb.clus_old = take(nc / g.cells_per_cl);
    s.sb.rank = nullptr;
    s.keys = (uint32_t *)take(m);
    s.stage = (double *)take(2 * m);
    s.err = take(1);
    return s;
}

int smi::seams::sort_exchange_impl(const Geom &g, int64_t n, const SortArrays &A, SortScratch &s, hipStream_t st)
{
    launch_fill_i32(s.sb.counts, g.ncells, 0, st);
    launch_fill_i32((int32_t *)s.sb.stay, (g.ncells + 3) / 4, 0, st);
    launch_fill_i32(s.err, 1, 0, st);
    if (n > 0) {
        if (g.nd == 3) hipLaunchKernelGGL(k_wrap<3>, dim3(nblk(n, 256)), dim3(256), 0, st, g, A.x, A.y, A.z, n, s.err);
        else hipLaunchKernelGGL(k_wrap<2>, dim3(nblk(n, 256)), dim3(256), 0, st, g, A.x, A.y, A.z, n, s.err);
        SEAM_CHECK();
        launch_bin_initial(g, PartPtrs{A.x, A.y, A.z, A.px, A.py, A.pz, A.w}, n, s.keys, st);
    }
    launch_sort(g, s.keys, n, s.sb, st, /*counted=*/false, /*in_cell_order=*/true, /*interleaved=*/false,
                /*tie=*/nullptr, /*placed=*/false);
    if (n > 0) {
        double *arrs[7] = {A.x, A.y, g.nd == 3 ? A.z : nullptr, A.px, A.py, A.pz, A.w};
        for (double *a : arrs) {
            if (!a) continue;
            hipLaunchKernelGGL(k_gather_perm<double>, dim3(nblk(n, 256)), dim3(256), 0, st,
                               (const int32_t *)s.sb.perm, (const double *)a, s.stage, n);
            SEAM_CHECK();
            if (hipMemcpyAsync(a, s.stage, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
                return op_fail(SMI_ERR_HIP, "device copy failed");
        }
        if (A.charge) {
            int16_t *st16 = (int16_t *)s.stage;
            hipLaunchKernelGGL(k_gather_perm<int16_t>, dim3(nblk(n, 256)), dim3(256), 0, st,
                               (const int32_t *)s.sb.perm, (const int16_t *)A.charge, st16, n);
            SEAM_CHECK();
            if (hipMemcpyAsync(A.charge, st16, n * sizeof(int16_t), hipMemcpyDeviceToDevice, st) != hipSuccess)
                return op_fail(SMI_ERR_HIP, "device copy failed");
        }
        if (A.key) {
            hipLaunchKernelGGL(k_keys_out, dim3(nblk(n, 256)), dim3(256), 0, st, g, (const int32_t *)s.sb.perm,
                               (const uint32_t *)s.keys, A.key, n);
            SEAM_CHECK();
        }
    }
    if (A.patch_count || A.bin_index) {
        hipLaunchKernelGGL(k_bins_out, dim3(nblk((int64_t)g.npatches * g.clpp, 256)), dim3(256), 0, st, g,
                           (const int32_t *)s.sb.cell_off, A.patch_count, A.bin_index);
        SEAM_CHECK();
    }
    int h = 0;
    hipError_t e = hipMemcpyAsync(&h, s.err, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return op_fail(SMI_ERR_HIP, hipGetErrorString(e));
    if (h) return op_fail(SMI_ERR_INVALID, "a particle lies farther than one box length outside the box");
    return SMI_OK;
}

int smi::seams::sort_exchange_args(int ndim, const int n_cells[3], const int patch_size[3], const double cell_length[3],
                       int64_t n, const SortArrays &A, Geom &g)
{
    if (!n_cells || !patch_size || !cell_length) return op_fail(SMI_ERR_INVALID, "null geometry");
    int rc = make_geom(ndim, 2, n_cells, patch_size, cell_length, g);
    if (rc) return rc;
    if (n < 0 || n >= (1ll << 31)) return op_fail(SMI_ERR_INVALID, "particle count out of range");
    if (n > 0 && (!A.x || !A.y || (ndim == 3 && !A.z) || !A.px || !A.py || !A.pz || !A.w))
        return op_fail(SMI_ERR_INVALID, "null particle array");
    return SMI_OK;
}

extern "C" {

int smi_current_deposition_3d(double *Jx, double *Jy, double *Jz, int Jx_size, int Jy_size, int Jz_size,
                              const double *particle_position_x, const double *particle_position_y,
                              const double *particle_position_z, const short *particle_charge,
                              const double *particle_weight, const int *bin_index, unsigned int x_dimension_bin_count,
                              unsigned int y_dimension_bin_count, unsigned int z_dimension_bin_count,
                              const double *invgf, const int *iold, const double *deltaold,
                              unsigned int number_of_particles, double inv_cell_volume, double dx_inv, double dy_inv,
                              double dz_inv, double dx_ov_dt, double dy_ov_dt, double dz_ov_dt, int i_domain_begin,
                              int j_domain_begin, int k_domain_begin, int nprimy, int nprimz, int not_spectral,
                              bool cell_sorting)
{
    (void)invgf;  // read by the 2D/AM projectors only
    SeamArgs a{};
    int rc = seam_args(a, bin_index, x_dimension_bin_count, y_dimension_bin_count, z_dimension_bin_count, iold,
                       deltaold, number_of_particles, inv_cell_volume, dx_inv, dy_inv, dz_inv, dx_ov_dt, dy_ov_dt,
                       dz_ov_dt, i_domain_begin, j_domain_begin, k_domain_begin, nprimy, nprimz, not_spectral,
                       cell_sorting);
    if (rc) return rc;
    if (!Jx || !Jy || !Jz) return op_fail(SMI_ERR_INVALID, "null current array");
    if (number_of_particles > 0 &&
        (!particle_position_x || !particle_position_y || !particle_position_z || !particle_charge || !particle_weight))
        return op_fail(SMI_ERR_INVALID, "null particle array");
    a.out[0] = Jx;
    a.out[1] = Jy;
    a.out[2] = Jz;
    a.size[0] = Jx_size;
    a.size[1] = Jy_size;
    a.size[2] = Jz_size;
    a.pos[0] = particle_position_x;
    a.pos[1] = particle_position_y;
    a.pos[2] = particle_position_z;
    a.charge = particle_charge;
    a.weight = particle_weight;
    return run_seam(k_seam_currents, a);
}

int smi_density_deposition_3d(double *rho, int rho_size, const double *particle_position_x,
                              const double *particle_position_y, const double *particle_position_z,
                              const short *particle_charge, const double *particle_weight, const int *bin_index,
                              unsigned int x_dimension_bin_count, unsigned int y_dimension_bin_count,
                              unsigned int z_dimension_bin_count, const double *invgf, const int *iold,
                              const double *deltaold, unsigned int number_of_particles, double inv_cell_volume,
                              double dx_inv, double dy_inv, double dz_inv, double dx_ov_dt, double dy_ov_dt,
                              double dz_ov_dt, int i_domain_begin, int j_domain_begin, int k_domain_begin, int nprimy,
                              int nprimz, int not_spectral, bool cell_sorting)
{
    (void)invgf;
    SeamArgs a{};
    int rc = seam_args(a, bin_index, x_dimension_bin_count, y_dimension_bin_count, z_dimension_bin_count, iold,
                       deltaold, number_of_particles, inv_cell_volume, dx_inv, dy_inv, dz_inv, dx_ov_dt, dy_ov_dt,
                       dz_ov_dt, i_domain_begin, j_domain_begin, k_domain_begin, nprimy, nprimz, not_spectral,
                       cell_sorting);
    if (rc) return rc;
    if (!rho) return op_fail(SMI_ERR_INVALID, "null density array");
    if (number_of_particles > 0 &&
        (!particle_position_x || !particle_position_y || !particle_position_z || !particle_charge || !particle_weight))
        return op_fail(SMI_ERR_INVALID, "null particle array");
    a.out[0] = rho;
    a.size[0] = rho_size;
    a.pos[0] = particle_position_x;
    a.pos[1] = particle_position_y;
    a.pos[2] = particle_position_z;
    a.charge = particle_charge;
    a.weight = particle_weight;
    return run_seam(k_seam_density, a);
}

int smi_density_3d(int order, double *rho, const int np[3], const int dom_begin[3], const double cell_length[3],
                   double charge, int64_t n, const double *x, const double *y, const double *z, const double *w,
                   void *stream)
{
    if (order != 2 && order != 4) return op_fail(SMI_ERR_INVALID, "order must be 2 or 4");
    if (n <= 0) return SMI_OK;
    if (!rho || !x || !y || !z || !w) return op_fail(SMI_ERR_INVALID, "null array");
    BasicArgs a{};
    a.rho = rho;
    double cell_volume = 1.0;
    for (int d = 0; d < 3; ++d) {
        a.np[d] = np[d];
        a.dom[d] = dom_begin[d];
        a.dinv[d] = 1.0 / cell_length[d];
        cell_volume *= cell_length[d];
    }
    a.cw_scale = (1. / cell_volume) * charge;
    a.n = n;
    a.pos[0] = x;
    a.pos[1] = y;
    a.pos[2] = z;
    a.w = w;
    return run_basic(3, order, a, (hipStream_t)stream);
}

int smi_density_2d(int order, double *rho, const int np[2], const int dom_begin[2], const double cell_length[2],
                   double charge, int64_t n, const double *x, const double *y, const double *w, void *stream)
{
    if (order != 2) return op_fail(SMI_ERR_INVALID, "2D order must be 2");
    if (n <= 0) return SMI_OK;
    if (!rho || !x || !y || !w) return op_fail(SMI_ERR_INVALID, "null array");
    BasicArgs a{};
    a.rho = rho;
    double cell_volume = 1.0;
    for (int d = 0; d < 2; ++d) {
        a.np[d] = np[d];
        a.dom[d] = dom_begin[d];
        a.dinv[d] = 1.0 / cell_length[d];
        cell_volume *= cell_length[d];
    }
    a.np[2] = 1;
    a.cw_scale = (1. / cell_volume) * charge;
    a.n = n;
    a.pos[0] = x;
    a.pos[1] = y;
    a.pos[2] = nullptr;
    a.w = w;
    return run_basic(2, 2, a, (hipStream_t)stream);
}

int smi_maxwell_ampere_2d(double *Ex, double *Ey, double *Ez, const double *Bx, const double *By, const double *Bz,
                          const double *Jx, const double *Jy, const double *Jz, const int np[2], double timestep,
                          const double cell_length[2], void *stream)
{
    if (np[0] < 2 || np[1] < 2) return op_fail(SMI_ERR_INVALID, "patch arrays too small");
    const int64_t tot = (int64_t)(np[0] + 1) * np[1] + (int64_t)np[0] * (np[1] + 1) + (int64_t)np[0] * np[1];
    hipLaunchKernelGGL(k_ma2d_patch, dim3(nblk(tot, 256)), dim3(256), 0, (hipStream_t)stream, Ex, Ey, Ez, Bx, By, Bz,
                       Jx, Jy, Jz, np[0], np[1], timestep, timestep / cell_length[0], timestep / cell_length[1]);
    SEAM_CHECK();
    return SMI_OK;
}

int smi_maxwell_faraday_2d(double *Bx, double *By, double *Bz, const double *Ex, const double *Ey, const double *Ez,
                           const int np[2], double timestep, const double cell_length[2], void *stream)
{
    if (np[0] < 2 || np[1] < 2) return op_fail(SMI_ERR_INVALID, "patch arrays too small");
    const int64_t tot = (int64_t)np[0] * (np[1] + 1) + (int64_t)(np[0] + 1) * np[1] + (int64_t)(np[0] + 1) * (np[1] + 1);
    hipLaunchKernelGGL(k_mf2d_patch, dim3(nblk(tot, 256)), dim3(256), 0, (hipStream_t)stream, Bx, By, Bz, Ex, Ey, Ez,
                       np[0], np[1], timestep / cell_length[0], timestep / cell_length[1]);
    SEAM_CHECK();
    return SMI_OK;
}

int smi_sum_j_ghosts(int ndim, int order, const int n_cells[3], const int patch_size[3], double *J, void *stream)
{
    if (!n_cells || !patch_size || !J) return op_fail(SMI_ERR_INVALID, "null argument");
    const double unit[3] = {1., 1., 1.};
    Geom g;
    int rc = make_geom(ndim, order, n_cells, patch_size, unit, g);
    if (rc) return rc;
    try {
        launch_ghost_sum(g, J, (hipStream_t)stream);
    } catch (const std::exception &e) {
        return op_fail(SMI_ERR_HIP, e.what());
    }
    return SMI_OK;
}

int smi_exchange_b(int ndim, int order, const int n_cells[3], const int patch_size[3], double *B, void *stream)
{
    if (!n_cells || !patch_size || !B) return op_fail(SMI_ERR_INVALID, "null argument");
    const double unit[3] = {1., 1., 1.};
    Geom g;
    int rc = make_geom(ndim, order, n_cells, patch_size, unit, g);
    if (rc) return rc;
    const PatchLayout L = patch_layout(g, kDualB);
    for (int d = 0; d < g.nd; ++d) {
        int64_t per = 0;
        for (int c = 0; c < 3; ++c) {
            if (c == d) continue;
            int64_t sz = 1;
            for (int e = 0; e < 3; ++e) sz *= e == d ? g.os : L.dims[c][e];
            per += sz;
        }
        const int64_t tot = per * g.npatches;
        hipLaunchKernelGGL(k_exchange_b, dim3(nblk(tot, 256)), dim3(256), 0, (hipStream_t)stream, g, L, d, B);
        SEAM_CHECK();
    }
    return SMI_OK;
}

int smi_sort_exchange(int ndim, const int n_cells[3], const int patch_size[3], const double cell_length[3], int64_t n,
                      double *x, double *y, double *z, double *px, double *py, double *pz, double *w, int16_t *charge,
                      int32_t *key, int32_t *patch_count, int32_t *bin_index, void *stream)
{
    const SortArrays A{x, y, z, px, py, pz, w, charge, key, patch_count, bin_index};
    Geom g;
    int rc = sort_exchange_args(ndim, n_cells, patch_size, cell_length, n, A, g);
    if (rc) return rc;
    int32_t *buf = nullptr;
    const int64_t words = sort_scratch_words(g, n);
    if (hipMalloc(&buf, words * 4) != hipSuccess) return op_fail(SMI_ERR_HIP, "hipMalloc of the sort scratch failed");
    SortScratch s = carve_sort_scratch(g, n, buf);
    try {
        rc = sort_exchange_impl(g, n, A, s, (hipStream_t)stream);
    } catch (const std::exception &e) {
        rc = op_fail(SMI_ERR_HIP, e.what());
    }
    (void)hipStreamSynchronize((hipStream_t)stream);
    (void)hipFree(buf);
    return rc;
}

}  // extern "C"
