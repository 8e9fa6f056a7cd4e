// fields.hip -- Yee field solver and diagnostics on the periodic global grid.
//
// The reference keeps one ghosted array per patch and synchronises the
// ghost planes after every update (SyncVectorPatch::sumRhoJ / exchangeB).
// With periodic boundaries every ghost value equals the owning patch's
// interior value, so one periodic array per component holds exactly the
// same numbers and each update below is a single coalesced stream over it.
// The arithmetic of each point follows MA_Solver3D_norm.cpp:18-115,
// MF_Solver3D_Yee.cpp:17-110 (and the 2D forms) term by term; B_m <- B
// (saveMagneticFields) and B_m <- (B + B_m)/2 (centerMagneticFields) are
// fused into the Faraday kernel.
#include "kernels.h"

namespace smi {

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct Idx3 {
    int i, j, k;
};

__device__ inline Idx3 decode(const Geom &g, int64_t t)
{
    Idx3 r;
    r.k = (int)(t % g.N[2]);
    r.j = (int)((t / g.N[2]) % g.N[1]);
    r.i = (int)(t / ((int64_t)g.N[2] * g.N[1]));
    return r;
}

__device__ inline int64_t lin(const Geom &g, int i, int j, int k)
{
    return ((int64_t)wrapi(i, g.N[0]) * g.N[1] + wrapi(j, g.N[1])) * g.N[2] + wrapi(k, g.N[2]);
}

// Maxwell-Ampere: E += -dt*J + dt/d x curl(B)  (all three components).
__global__ void k_ampere3d(Geom g, FieldSet f, double dtx, double dty, double dtz)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const Idx3 q = decode(g, t);
    const double dt = g.dt;
    const double *Bx = f.B[0], *By = f.B[1], *Bz = f.B[2];
    // Ex(d,p,p) += -dt*Jx + dt/dy*(Bz[j+1]-Bz[j]) - dt/dz*(By[k+1]-By[k])
    f.E[0][t] += -dt * f.J[0][t] + dty * (Bz[lin(g, q.i, q.j + 1, q.k)] - Bz[t]) -
                 dtz * (By[lin(g, q.i, q.j, q.k + 1)] - By[t]);
    // Ey(p,d,p) += -dt*Jy - dt/dx*(Bz[i+1]-Bz[i]) + dt/dz*(Bx[k+1]-Bx[k])
    f.E[1][t] += -dt * f.J[1][t] - dtx * (Bz[lin(g, q.i + 1, q.j, q.k)] - Bz[t]) +
                 dtz * (Bx[lin(g, q.i, q.j, q.k + 1)] - Bx[t]);
    // Ez(p,p,d) += -dt*Jz + dt/dx*(By[i+1]-By[i]) - dt/dy*(Bx[j+1]-Bx[j])
    f.E[2][t] += -dt * f.J[2][t] + dtx * (By[lin(g, q.i + 1, q.j, q.k)] - By[t]) -
                 dty * (Bx[lin(g, q.i, q.j + 1, q.k)] - Bx[t]);
}

// Maxwell-Faraday + B_m centring:  Bold = B; B += -dt/d x curl(E);
// B_m = (B + Bold) * 0.5.
__global__ void k_faraday3d(Geom g, FieldSet f, double dtx, double dty, double dtz)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const Idx3 q = decode(g, t);
    const double *Ex = f.E[0], *Ey = f.E[1], *Ez = f.E[2];
    // Bx(p,d,d) += -dt/dy*(Ez[j]-Ez[j-1]) + dt/dz*(Ey[k]-Ey[k-1])
    const double bx0 = f.B[0][t];
    const double bx = bx0 + (-dty * (Ez[t] - Ez[lin(g, q.i, q.j - 1, q.k)]) +
                             dtz * (Ey[t] - Ey[lin(g, q.i, q.j, q.k - 1)]));
    // By(d,p,d) += -dt/dz*(Ex[k]-Ex[k-1]) + dt/dx*(Ez[i]-Ez[i-1])
    const double by0 = f.B[1][t];
    const double by = by0 + (-dtz * (Ex[t] - Ex[lin(g, q.i, q.j, q.k - 1)]) +
                             dtx * (Ez[t] - Ez[lin(g, q.i - 1, q.j, q.k)]));
    // Bz(d,d,p) += -dt/dx*(Ey[i]-Ey[i-1]) + dt/dy*(Ex[j]-Ex[j-1])
    const double bz0 = f.B[2][t];
    const double bz = bz0 + (-dtx * (Ey[t] - Ey[lin(g, q.i - 1, q.j, q.k)]) +
                             dty * (Ex[t] - Ex[lin(g, q.i, q.j - 1, q.k)]));
    f.B[0][t] = bx;
    f.B[1][t] = by;
    f.B[2][t] = bz;
    f.Bm[0][t] = (bx + bx0) * 0.5;
    f.Bm[1][t] = (by + by0) * 0.5;
    f.Bm[2][t] = (bz + bz0) * 0.5;
}

// 2D forms (MA_Solver2D_norm.cpp:18-, MF_Solver2D_Yee.cpp:18-).
__global__ void k_ampere2d(Geom g, FieldSet f, double dtx, double dty)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const Idx3 q = decode(g, t);
    const double dt = g.dt;
    const double *Bx = f.B[0], *By = f.B[1], *Bz = f.B[2];
    f.E[0][t] += -dt * f.J[0][t] + dty * (Bz[lin(g, q.i, q.j + 1, 0)] - Bz[t]);
    f.E[1][t] += -dt * f.J[1][t] - dtx * (Bz[lin(g, q.i + 1, q.j, 0)] - Bz[t]);
    f.E[2][t] += -dt * f.J[2][t] + dtx * (By[lin(g, q.i + 1, q.j, 0)] - By[t]) -
                 dty * (Bx[lin(g, q.i, q.j + 1, 0)] - Bx[t]);
}

__global__ void k_faraday2d(Geom g, FieldSet f, double dtx, double dty)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const Idx3 q = decode(g, t);
    const double *Ex = f.E[0], *Ey = f.E[1], *Ez = f.E[2];
    const double bx0 = f.B[0][t], by0 = f.B[1][t], bz0 = f.B[2][t];
    const double bx = bx0 - dty * (Ez[t] - Ez[lin(g, q.i, q.j - 1, 0)]);
    const double by = by0 + dtx * (Ez[t] - Ez[lin(g, q.i - 1, q.j, 0)]);
    const double bz = bz0 + (dty * (Ex[t] - Ex[lin(g, q.i, q.j - 1, 0)]) -
                             dtx * (Ey[t] - Ey[lin(g, q.i - 1, q.j, 0)]));
    f.B[0][t] = bx;
    f.B[1][t] = by;
    f.B[2][t] = bz;
    f.Bm[0][t] = (bx + bx0) * 0.5;
    f.Bm[1][t] = (by + by0) * 0.5;
    f.Bm[2][t] = (bz + bz0) * 0.5;
}

__global__ void k_fill(double *p, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ inline double block_sum(double v)
{
    __shared__ double part[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    double s = 0.;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += part[w];
    __syncthreads();
    return s;
}

struct FieldList {
    const double *f[6];
    int dual[6];  // staggering bits, axis d at bit d
    int nf;
};

// Sum of squares with the point multiplicities of Field3D::norm2 over the
// patches' istart/bufsize ranges (ElectroMagn3D.cpp:185-226,
// Field3D.cpp:230-250): along each axis of the simulation those ranges cover
// N+1 primal points (global 0..N, N being the periodic image of 0) and N+2
// dual points (0..N+1, images of 0 and 1), so in the periodic global array
// primal index 0 and dual indices 0 and 1 count twice.
__global__ void k_sumsq(Geom g, FieldList L, double *partials)
{
    double s = 0.;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.ncells;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int idx[3] = {(int)(t / ((int64_t)g.N[1] * g.N[2])), (int)((t / g.N[2]) % g.N[1]), (int)(t % g.N[2])};
        for (int c = 0; c < L.nf; ++c) {
            double m = 1.;
            for (int d = 0; d < g.nd; ++d) {
                const bool dual = (L.dual[c] >> d) & 1;
                if (idx[d] == 0 || (dual && idx[d] == 1)) m *= 2.;
            }
            const double v = L.f[c][t];
            s += m * (v * v);
        }
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sum of the weights w[0..n) (FP64 array w, or FP32 array wf).
__global__ void k_wsum(const double *__restrict__ w, const float *__restrict__ wf, int64_t n, double *partials)
{
    double s = 0.;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
        s += w ? w[j] : (double)wf[j];
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void k_ukin(PartPtrs A, const int32_t *__restrict__ perm, int64_t n, double *partials)
{
    double s = 0.;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const int i = perm[j];
        const double ux = A.px[i], uy = A.py[i], uz = A.pz[i];
        s += A.w[i] * (sqrt(1 + ux * ux + uy * uy + uz * uz) - 1.0);
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Charge density, shape of the interpolation order (Projector3D2Order::
// basic type 0, Projector3D2Order.cpp:526-621), FP64 atomics, periodic.
template <int ND, int ORDER>
__global__ void k_rho(Geom g, SpeciesK sp, PartPtrs A, const int32_t *__restrict__ perm, int64_t n,
                      double *__restrict__ rho)
{
    constexpr int H = ORDER / 2, W = 2 * H + 1;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int i = perm[j];
    const double x[3] = {A.x[i], A.y[i], ND == 3 ? A.z[i] : 0.};
    double S[3][5];
    int o[3] = {0, 0, 0};
    for (int d = 0; d < ND; ++d) {
        const double pn = x[d] * g.dinv[d];
        if (!(pn > -4.0 && pn < g.N[d] + 4.0)) return;  // outside the box: not deposited
        const int ip = (int)round(pn);
        spline<ORDER>(pn - (double)ip, S[d]);
        o[d] = ip - H;
    }
    const double cw = sp.cw_scale * A.w[i];
    for (int a = 0; a < W; ++a)
        for (int b = 0; b < W; ++b) {
            if constexpr (ND == 3) {
                for (int c = 0; c < W; ++c)
                    unsafeAtomicAdd(&rho[lin(g, o[0] + a, o[1] + b, o[2] + c)], cw * S[0][a] * S[1][b] * S[2][c]);
            } else {
                unsafeAtomicAdd(&rho[lin(g, o[0] + a, o[1] + b, 0)], cw * S[0][a] * S[1][b]);
            }
        }
}

// Global key of cell c (periodic indices in [0, N)): patch, 4^ND cluster,
// cell in the cluster (Cluster3D::CellKey, nvidiaParticles.cu:639-660).
template <int ND>
__device__ inline uint32_t global_cell_key(const Geom &g, const int *c)
{
    int pc[3] = {0, 0, 0}, l[3] = {0, 0, 0};
    for (int d = 0; d < ND; ++d) {
        pc[d] = c[d] / g.n[d];
        l[d] = c[d] - pc[d] * g.n[d];
    }
    if constexpr (ND == 3) {
        const int p = (pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2];
        const int cl = ((l[0] >> 2) * g.ncl[1] + (l[1] >> 2)) * g.ncl[2] + (l[2] >> 2);
        return (uint32_t)((int64_t)p * g.cpp + cl * 64 + 16 * (l[0] & 3) + 4 * (l[1] & 3) + (l[2] & 3));
    } else {
        const int p = pc[0] * g.npatch[1] + pc[1];
        const int cl = (l[0] >> 2) * g.ncl[1] + (l[1] >> 2);
        return (uint32_t)((int64_t)p * g.cpp + cl * 16 + 4 * (l[0] & 3) + (l[1] & 3));
    }
}

// Charge density by gather, for the deterministic deposit mode: one thread
// per grid node adds to rho[node] the contributions (same shape and product
// order as k_rho) of the particles of the (ORDER+2)^ND cells whose stencils
// can reach the node, cells in x, y, z order and particles in their sorted
// order inside each cell.  The summation order is fixed, so the density of
// a state (and the Poisson field of a namelist start) is reproducible bit
// for bit, which the atomic scatter is not.  cell_off: the species' sorted
// cell ranges (cell-major ranks; stored in the cluster layout lay).
template <int ND, int ORDER>
__global__ void k_rho_gather(Geom g, SpeciesK sp, PartPtrs A, const int32_t *__restrict__ perm,
                             const int32_t *__restrict__ cell_off, int lay, double *__restrict__ rho)
{
    constexpr int H = ORDER / 2, W = 2 * H + 1, R = 2 * H + 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const Idx3 q = decode(g, t);
    const int node[3] = {q.i, q.j, q.k};
    const double *X[3] = {A.x, A.y, A.z};
    const int cpc = g.cells_per_cl;
    double acc = 0.;
    for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b)
            for (int e = 0; e < (ND == 3 ? R : 1); ++e) {
                // source cell: unwrapped index cu, periodic image cw
                const int cu[3] = {node[0] - H - 1 + a, node[1] - H - 1 + b, ND == 3 ? node[2] - H - 1 + e : 0};
                int cw[3] = {0, 0, 0};
                for (int d = 0; d < ND; ++d) cw[d] = wrapi(cu[d], g.N[d]);
                const uint32_t key = global_cell_key<ND>(g, cw);
                const int r0 = cell_off[key], r1 = cell_off[key + 1];
                if (r0 == r1) continue;
                int cs = 0, cn = 0;
                if (lay) {
                    const uint32_t cb = key - key % (uint32_t)cpc;
                    cs = cell_off[cb];
                    cn = cell_off[cb + cpc] - cs;
                }
                for (int r = r0; r < r1; ++r) {
                    const int i = perm[lay ? cs + layout_slot(lay, r - cs, cn) : r];
                    double S[3][5];
                    int o[3] = {0, 0, 0};
                    bool reach = true;
                    for (int d = 0; d < ND && reach; ++d) {
                        const double pn = X[d][i] * g.dinv[d];
                        const int ip = (int)round(pn);
                        o[d] = node[d] - (ip + (cu[d] - cw[d]) - H);
                        reach = o[d] >= 0 && o[d] < W;
                        if (reach) spline<ORDER>(pn - (double)ip, S[d]);
                    }
                    if (!reach) continue;
                    const double cwp = sp.cw_scale * A.w[i];
                    if constexpr (ND == 3) acc += cwp * S[0][o[0]] * S[1][o[1]] * S[2][o[2]];
                    else acc += cwp * S[0][o[0]] * S[1][o[1]];
                }
            }
    rho[t] += acc;
}

// residual = div(E) - rho on primal nodes (dual g+1 sits at (g+1/2)dx).
__global__ void k_gauss(Geom g, FieldSet f, double *__restrict__ r)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncells) return;
    const Idx3 q = decode(g, t);
    double div = (f.E[0][lin(g, q.i + 1, q.j, q.k)] - f.E[0][t]) / g.dx[0];
    if (g.nd > 1) div += (f.E[1][lin(g, q.i, q.j + 1, q.k)] - f.E[1][t]) / g.dx[1];
    if (g.nd > 2) div += (f.E[2][lin(g, q.i, q.j, q.k + 1)] - f.E[2][t]) / g.dx[2];
    r[t] = div - r[t];
}

}  // namespace

void launch_maxwell(const Geom &g, const FieldSet &f, hipStream_t st)
{
    const double dtx = g.dt / g.dx[0], dty = g.dt / g.dx[1], dtz = g.dt / g.dx[2];
    const unsigned nb = nblk(g.ncells, 256);
    if (g.nd == 3) {
        hipLaunchKernelGGL(k_ampere3d, dim3(nb), dim3(256), 0, st, g, f, dtx, dty, dtz);
        SMI_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_faraday3d, dim3(nb), dim3(256), 0, st, g, f, dtx, dty, dtz);
        SMI_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(k_ampere2d, dim3(nb), dim3(256), 0, st, g, f, dtx, dty);
        SMI_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_faraday2d, dim3(nb), dim3(256), 0, st, g, f, dtx, dty);
        SMI_LAUNCH_CHECK();
    }
}

void launch_fill(double *p, int64_t n, double v, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
    SMI_LAUNCH_CHECK();
}

void launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
    SMI_LAUNCH_CHECK();
}

void launch_sumsq(const Geom &g, const double *const *fields, const int *fids, int nf, double *partials,
                  int nblocks, hipStream_t st)
{
    // staggering of Ex Ey Ez Bx By Bz Bx_m By_m Bz_m Jx Jy Jz (bit d: dual along d)
    static const int kDualBits[12] = {1, 2, 4, 6, 5, 3, 6, 5, 3, 1, 2, 4};
    if (nf < 1 || nf > 6) throw Error(-1, "sumsq takes 1 to 6 fields");
    FieldList L;
    L.nf = nf;
    for (int c = 0; c < nf; ++c) {
        if (fids[c] < 0 || fids[c] >= 12) throw Error(-1, "bad field id");
        L.f[c] = fields[c];
        L.dual[c] = kDualBits[fids[c]];
    }
    hipLaunchKernelGGL(k_sumsq, dim3(nblocks), dim3(256), 0, st, g, L, partials);
    SMI_LAUNCH_CHECK();
}

void launch_wsum(const double *w, const float *wf, int64_t n, double *partials, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_wsum, dim3(nblocks), dim3(256), 0, st, w, wf, n, partials);
    SMI_LAUNCH_CHECK();
}

void launch_ukin(const PartPtrs &A, const int32_t *perm, int64_t n, double *partials, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_ukin, dim3(nblocks), dim3(256), 0, st, A, perm, n, partials);
    SMI_LAUNCH_CHECK();
}

void launch_rho(const Geom &g, const SpeciesK &sp, const PartPtrs &A, const int32_t *perm, int64_t n, double *rho,
                hipStream_t st)
{
    if (n == 0) return;
    if (g.nd == 3 && g.order == 2)
        hipLaunchKernelGGL((k_rho<3, 2>), dim3(nblk(n, 256)), dim3(256), 0, st, g, sp, A, perm, n, rho);
    else if (g.nd == 3 && g.order == 4)
        hipLaunchKernelGGL((k_rho<3, 4>), dim3(nblk(n, 256)), dim3(256), 0, st, g, sp, A, perm, n, rho);
    else
        hipLaunchKernelGGL((k_rho<2, 2>), dim3(nblk(n, 256)), dim3(256), 0, st, g, sp, A, perm, n, rho);
    SMI_LAUNCH_CHECK();
}

void launch_rho_gather(const Geom &g, const SpeciesK &sp, const PartPtrs &A, const int32_t *perm,
                       const int32_t *cell_off, int lay, double *rho, hipStream_t st)
{
    const dim3 grid(nblk(g.ncells, 256));
    if (g.nd == 3 && g.order == 2)
        hipLaunchKernelGGL((k_rho_gather<3, 2>), grid, dim3(256), 0, st, g, sp, A, perm, cell_off, lay, rho);
    else if (g.nd == 3 && g.order == 4)
        hipLaunchKernelGGL((k_rho_gather<3, 4>), grid, dim3(256), 0, st, g, sp, A, perm, cell_off, lay, rho);
    else
        hipLaunchKernelGGL((k_rho_gather<2, 2>), grid, dim3(256), 0, st, g, sp, A, perm, cell_off, lay, rho);
    SMI_LAUNCH_CHECK();
}

void launch_gauss(const Geom &g, const FieldSet &f, double *r, hipStream_t st)
{
    hipLaunchKernelGGL(k_gauss, dim3(nblk(g.ncells, 256)), dim3(256), 0, st, g, f, r);
    SMI_LAUNCH_CHECK();
}

}  // namespace smi
