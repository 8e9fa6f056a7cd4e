// operators.hip -- per-operator entry points in the reference's per-patch
// memory layout (ghosted Field3D/Field2D arrays, SoA particles and the
// SmileiMPI dynamics_* scratch with stride n).  A reference build that keeps
// its own orchestration calls these from its operator classes in place of
// the OpenMP-target loops and of currentDeposition3DOnDevice
// (src/Projector/Projector3D2OrderGPU.cpp:55-88).
#include "../../include/smilei_amd.h"
#include "kernels.h"
#include "pic_device.h"

using namespace smi;
using namespace smi::dev;

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

struct Ints3 {
    int v[3];
};
struct Dbl3 {
    double v[3];
};

// Interpolator3D{2,4}Order::fieldsWrapper, one thread per particle.
// (launched with 256 threads: the bound leaves the compiler 256 registers
// instead of the 128 of a 1024-thread default, so nothing spills)
template <int ORDER>
__global__ void __launch_bounds__(256, 2) k_interp3d(const double *Ex, const double *Ey, const double *Ez, const double *Bx, const double *By,
                           const double *Bz, Ints3 np, Ints3 dom, Dbl3 dinv, int64_t n, const double *x,
                           const double *y, const double *z, double *Ep, double *Bp, int32_t *iold, double *dold)
{
    constexpr int H = ORDER / 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pos[3] = {x[i], y[i], z[i]};
    double cp[3][5], cd[3][5], dp[3];
    int ip[3], id[3];
    for (int d = 0; d < 3; ++d) {
        const double pn = pos[d] * dinv.v[d];
        ip[d] = (int)round(pn);
        dp[d] = pn - (double)ip[d];
        spline<ORDER>(dp[d], cp[d]);
        id[d] = (int)round(pn + 0.5);
        spline<ORDER>(pn - (double)id[d] + 0.5, cd[d]);
        ip[d] -= dom.v[d];
        id[d] -= dom.v[d];
    }
    bool ok = true;
    for (int d = 0; d < 3; ++d)
        ok &= ip[d] >= H && ip[d] + H <= np.v[d] - 1 && id[d] >= H && id[d] + H <= np.v[d];
    if (!ok) {  // position outside the patch: no memory access, NaN result
        for (int c = 0; c < 3; ++c) Ep[c * n + i] = Bp[c * n + i] = __builtin_nan("");
        return;
    }
    const int ny = np.v[1], nz = np.v[2];
    Ep[i] = gather3s<H>(cd[0], cp[1], cp[2], Ex, id[0], ip[1], ip[2], ny, nz);
    Ep[n + i] = gather3s<H>(cp[0], cd[1], cp[2], Ey, ip[0], id[1], ip[2], ny + 1, nz);
    Ep[2 * n + i] = gather3s<H>(cp[0], cp[1], cd[2], Ez, ip[0], ip[1], id[2], ny, nz + 1);
    Bp[i] = gather3s<H>(cp[0], cd[1], cd[2], Bx, ip[0], id[1], id[2], ny + 1, nz + 1);
    Bp[n + i] = gather3s<H>(cd[0], cp[1], cd[2], By, id[0], ip[1], id[2], ny, nz + 1);
    Bp[2 * n + i] = gather3s<H>(cd[0], cd[1], cp[2], Bz, id[0], id[1], ip[2], ny + 1, nz);
    for (int d = 0; d < 3; ++d) {
        iold[d * n + i] = ip[d];
        dold[d * n + i] = dp[d];
    }
}

// Interpolator2D2Order::fieldsWrapper.
__global__ void k_interp2d(const double *Ex, const double *Ey, const double *Ez, const double *Bx, const double *By,
                           const double *Bz, Ints3 np, Ints3 dom, Dbl3 dinv, int64_t n, const double *x,
                           const double *y, double *Ep, double *Bp, int32_t *iold, double *dold)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pos[2] = {x[i], y[i]};
    double cp[2][5], cd[2][5], dp[2];
    int ip[2], id[2];
    for (int d = 0; d < 2; ++d) {
        const double pn = pos[d] * dinv.v[d];
        ip[d] = (int)round(pn);
        dp[d] = pn - (double)ip[d];
        spline2(dp[d], cp[d]);
        id[d] = (int)round(pn + 0.5);
        spline2(pn - (double)id[d] + 0.5, cd[d]);
        ip[d] -= dom.v[d];
        id[d] -= dom.v[d];
    }
    bool ok = true;
    for (int d = 0; d < 2; ++d) ok &= ip[d] >= 1 && ip[d] + 1 <= np.v[d] - 1 && id[d] >= 1 && id[d] + 1 <= np.v[d];
    if (!ok) {
        for (int c = 0; c < 3; ++c) Ep[c * n + i] = Bp[c * n + i] = __builtin_nan("");
        return;
    }
    const int ny = np.v[1];
    Ep[i] = gather2s<1>(cd[0], cp[1], Ex, id[0], ip[1], ny);
    Ep[n + i] = gather2s<1>(cp[0], cd[1], Ey, ip[0], id[1], ny + 1);
    Ep[2 * n + i] = gather2s<1>(cp[0], cp[1], Ez, ip[0], ip[1], ny);
    Bp[i] = gather2s<1>(cp[0], cd[1], Bx, ip[0], id[1], ny + 1);
    Bp[n + i] = gather2s<1>(cd[0], cp[1], By, id[0], ip[1], ny);
    Bp[2 * n + i] = gather2s<1>(cd[0], cd[1], Bz, id[0], id[1], ny + 1);
    for (int d = 0; d < 2; ++d) {
        iold[d * n + i] = ip[d];
        dold[d * n + i] = dp[d];
    }
}

// PusherBoris::operator().
__global__ void k_boris(int ndim, double qmh, double dt, int64_t n, double *x, double *y, double *z, double *px,
                        double *py, double *pz, const double *Ep, const double *Bp, double *invgf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double E[3] = {Ep[i], Ep[n + i], Ep[2 * n + i]};
    const double B[3] = {Bp[i], Bp[n + i], Bp[2 * n + i]};
    double u[3] = {px[i], py[i], pz[i]}, dxp[3], g;
    boris(qmh, dt, E, B, u, dxp, g);
    invgf[i] = g;
    px[i] = u[0];
    py[i] = u[1];
    pz[i] = u[2];
    x[i] += dxp[0];
    if (ndim > 1) {
        y[i] += dxp[1];
        if (ndim > 2) z[i] += dxp[2];
    }
}

struct Proj3Args {
    double *J[3];
    int dims[3][3];
    Ints3 dom;
    Dbl3 dinv, cr_scale;  // cr = cw * dx/dt
    double cw_scale;
    int64_t n;
    const double *x[3];
    const double *w;
    const int32_t *iold;
    const double *dold;
};

// Returns false (and touches no J point) when the particle's stencil would
// leave the patch arrays or it moved by more than one cell.
template <int ORDER>
__device__ inline bool proj3_particle_shapes(const Proj3Args &a, int64_t i, double (&S0)[3][7], double (&S1)[3][7],
                                             double (&DS)[3][7], int *io)
{
    constexpr int H = ORDER / 2;
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
        io[d] = a.iold[d * a.n + i];
        const double pn = a.x[d][i] * a.dinv.v[d];
        const int sh = (int)round(pn) - (io[d] + a.dom.v[d]);
        ok &= io[d] - (H + 1) >= 0 && io[d] + H + 1 <= a.dims[(d + 1) % 3][d] - 1 && sh >= -1 && sh <= 1;
        esirkepov_shapes<ORDER>(a.dold[d * a.n + i], pn, io[d] + a.dom.v[d], S0[d], S1[d], DS[d]);
    }
    return ok;
}

// Projector3D{2,4}Order::currents with FP64 atomics, one thread per particle.
// (launched with 128 threads; register budget as k_interp3d)
template <int ORDER>
__global__ void __launch_bounds__(128, 2) k_proj3d_atomic(Proj3Args a)
{
    constexpr int H = ORDER / 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double S0[3][7], S1[3][7], DS[3][7];
    int io[3];
    if (!proj3_particle_shapes<ORDER>(a, i, S0, S1, DS, io)) return;
    const double cw = a.cw_scale * a.w[i];
    const double cr[3] = {cw * a.cr_scale.v[0], cw * a.cr_scale.v[1], cw * a.cr_scale.v[2]};
    esirkepov3d<ORDER>(S0, DS, cr, [&](int c, int u, int v, int k, double val) {
        if (val == 0.) return;
        const int *dm = a.dims[c];
        const int64_t idx = ((int64_t)(io[0] - (H + 1) + u) * dm[1] + (io[1] - (H + 1) + v)) * dm[2] +
                            (io[2] - (H + 1) + k);
        unsafeAtomicAdd(&a.J[c][idx], val);
    });
}

// Same contributions in particle order: one wavefront, lanes over
// (component, transverse column), ordered L2 atomics and a barrier per
// particle; every J point sums in index order like the CPU operator.
template <int ORDER>
__global__ void __launch_bounds__(64) k_proj3d_ordered(Proj3Args a)
{
    constexpr int H = ORDER / 2, S = 2 * H + 3;
    for (int64_t i = 0; i < a.n; ++i) {
        double S0[3][7], S1[3][7], DS[3][7];
        int io[3];
        if (!proj3_particle_shapes<ORDER>(a, i, S0, S1, DS, io)) continue;  // uniform across the wavefront
        const double cw = a.cw_scale * a.w[i];
        const double cr[3] = {cw * a.cr_scale.v[0], cw * a.cr_scale.v[1], cw * a.cr_scale.v[2]};
        for (int task = threadIdx.x; task < 3 * S * S; task += 64) {
            const int c = task / (S * S), col = task % (S * S);
            const int u = col / S, v = col % S;
            const int *dm = a.dims[c];
            double W;
            if (c == 0) W = weight3(S0[1], DS[1], S0[2], DS[2], u, v);
            else if (c == 1) W = weight3(S0[2], DS[2], S0[0], DS[0], v, u);
            else W = weight3(S0[0], DS[0], S0[1], DS[1], u, v);
            const double t = cr[c] * W;
            double acc = 0.;
            for (int r = 1; r < S; ++r) {
                if constexpr (ORDER == 2) acc -= DS[c][r - 1] * t;
                else acc -= cr[c] * DS[c][r - 1] * W;
                int ii, jj, kk;
                if (c == 0) { ii = r; jj = u; kk = v; }
                else if (c == 1) { ii = u; jj = r; kk = v; }
                else { ii = u; jj = v; kk = r; }
                const int64_t idx = ((int64_t)(io[0] - (H + 1) + ii) * dm[1] + (io[1] - (H + 1) + jj)) * dm[2] +
                                    (io[2] - (H + 1) + kk);
                unsafeAtomicAdd(&a.J[c][idx], acc);
            }
        }
        __syncthreads();
    }
}

struct Proj2Args {
    double *J[3];
    int ny[3];  // second extent of Jx, Jy, Jz
    int nx;     // primal extent along x
    Ints3 dom;
    Dbl3 dinv, cr_scale;
    double cw_scale;
    int64_t n;
    const double *x, *y, *pz, *w, *invgf;
    const int32_t *iold;
    const double *dold;
};

__device__ inline bool proj2_particle(const Proj2Args &a, int64_t i, double (&S0)[3][7], double (&S1)[3][7],
                                      double (&DS)[3][7], int *io, double &crx, double &cry, double &crz)
{
    const double pos[2] = {a.x[i], a.y[i]};
    bool ok = true;
    for (int d = 0; d < 2; ++d) {
        io[d] = a.iold[d * a.n + i];
        const double pn = pos[d] * a.dinv.v[d];
        const int sh = (int)round(pn) - (io[d] + a.dom.v[d]);
        // Jz (p,p) has the primal extents: ny[2] along y, nprim_x along x
        const int ext = d == 0 ? a.nx : a.ny[2];
        ok &= io[d] - 2 >= 0 && io[d] + 2 <= ext - 1 && sh >= -1 && sh <= 1;
        esirkepov_shapes<2>(a.dold[d * a.n + i], pn, io[d] + a.dom.v[d], S0[d], S1[d], DS[d]);
    }
    const double cw = a.cw_scale * a.w[i];
    crx = cw * a.cr_scale.v[0];
    cry = cw * a.cr_scale.v[1];
    crz = cw * kThird * a.pz[i] * a.invgf[i];
    return ok;
}

__global__ void k_proj2d_atomic(Proj2Args a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double S0[3][7], S1[3][7], DS[3][7], crx, cry, crz;
    int io[2];
    if (!proj2_particle(a, i, S0, S1, DS, io, crx, cry, crz)) return;
    esirkepov2d(S0, S1, DS, crx, cry, crz, [&](int c, int u, int v, double val) {
        if (val == 0.) return;
        unsafeAtomicAdd(&a.J[c][(int64_t)(io[0] - 2 + u) * a.ny[c] + (io[1] - 2 + v)], val);
    });
}

__global__ void __launch_bounds__(64) k_proj2d_ordered(Proj2Args a)
{
    const int lane = threadIdx.x;
    for (int64_t i = 0; i < a.n; ++i) {
        double S0[3][7], S1[3][7], DS[3][7], crx, cry, crz;
        int io[2];
        if (!proj2_particle(a, i, S0, S1, DS, io, crx, cry, crz)) continue;  // uniform across the wavefront
        if (lane < 5) {
            const int jj = lane;
            double acc = 0.;
            for (int ii = 1; ii < 5; ++ii) {
                if (jj == 0) acc -= crx * DS[0][ii - 1] * (0.5 * DS[1][0]);
                else acc -= crx * DS[0][ii - 1] * (S0[1][jj] + 0.5 * DS[1][jj]);
                unsafeAtomicAdd(&a.J[0][(int64_t)(io[0] - 2 + ii) * a.ny[0] + (io[1] - 2 + jj)], acc);
            }
        } else if (lane < 10) {
            const int ii = lane - 5;
            const double ty = S0[0][ii] + 0.5 * DS[0][ii];
            double acc = 0.;
            for (int jj = 1; jj < 5; ++jj) {
                acc -= cry * DS[1][jj - 1] * ty;
                unsafeAtomicAdd(&a.J[1][(int64_t)(io[0] - 2 + ii) * a.ny[1] + (io[1] - 2 + jj)], acc);
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
            unsafeAtomicAdd(&a.J[2][(int64_t)(io[0] - 2 + ii) * a.ny[2] + (io[1] - 2 + jj)], v);
        }
        __syncthreads();
    }
}

// MA_Solver3D_norm on one patch array: each component over its own extent.
__global__ void k_ma3d_patch(double *Ex, double *Ey, double *Ez, const double *Bx, const double *By,
                             const double *Bz, const double *Jx, const double *Jy, const double *Jz, Ints3 np,
                             double dt, double dtx, double dty, double dtz)
{
    const int64_t nxp = np.v[0], nyp = np.v[1], nzp = np.v[2], nxd = nxp + 1, nyd = nyp + 1, nzd = nzp + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nEx = nxd * nyp * nzp, nEy = nxp * nyd * nzp, nEz = nxp * nyp * nzd;
    if (t < nEx) {
        const int64_t i = t / (nyp * nzp), j = (t / nzp) % nyp, k = t % nzp;
        Ex[t] += -dt * Jx[t] + dty * (Bz[(i * nyd + j + 1) * nzp + k] - Bz[(i * nyd + j) * nzp + k]) -
                 dtz * (By[(i * nyp + j) * nzd + k + 1] - By[(i * nyp + j) * nzd + k]);
    } else if (t < nEx + nEy) {
        const int64_t e = t - nEx;
        const int64_t i = e / (nyd * nzp), j = (e / nzp) % nyd, k = e % nzp;
        Ey[e] += -dt * Jy[e] - dtx * (Bz[((i + 1) * nyd + j) * nzp + k] - Bz[(i * nyd + j) * nzp + k]) +
                 dtz * (Bx[(i * nyd + j) * nzd + k + 1] - Bx[(i * nyd + j) * nzd + k]);
    } else if (t < nEx + nEy + nEz) {
        const int64_t e = t - nEx - nEy;
        const int64_t i = e / (nyp * nzd), j = (e / nzd) % nyp, k = e % nzd;
        Ez[e] += -dt * Jz[e] + dtx * (By[((i + 1) * nyp + j) * nzd + k] - By[(i * nyp + j) * nzd + k]) -
                 dty * (Bx[(i * nyd + j + 1) * nzd + k] - Bx[(i * nyd + j) * nzd + k]);
    }
}

// MF_Solver3D_Yee on one patch array: interior of the dual axes only.
__global__ void k_mf3d_patch(double *Bx, double *By, double *Bz, const double *Ex, const double *Ey,
                             const double *Ez, Ints3 np, double dtx, double dty, double dtz)
{
    const int64_t nxp = np.v[0], nyp = np.v[1], nzp = np.v[2], nxd = nxp + 1, nyd = nyp + 1, nzd = nzp + 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nBx = nxp * nyd * nzd, nBy = nxd * nyp * nzd, nBz = nxd * nyd * nzp;
    if (t < nBx) {
        const int64_t i = t / (nyd * nzd), j = (t / nzd) % nyd, k = t % nzd;
        if (j < 1 || j >= nyd - 1 || k < 1 || k >= nzd - 1) return;
        Bx[t] += -dty * (Ez[(i * nyp + j) * nzd + k] - Ez[(i * nyp + j - 1) * nzd + k]) +
                 dtz * (Ey[(i * nyd + j) * nzp + k] - Ey[(i * nyd + j) * nzp + k - 1]);
    } else if (t < nBx + nBy) {
        const int64_t e = t - nBx;
        const int64_t i = e / (nyp * nzd), j = (e / nzd) % nyp, k = e % nzd;
        if (i < 1 || i >= nxd - 1 || k < 1 || k >= nzd - 1) return;
        By[e] += -dtz * (Ex[(i * nyp + j) * nzp + k] - Ex[(i * nyp + j) * nzp + k - 1]) +
                 dtx * (Ez[(i * nyp + j) * nzd + k] - Ez[((i - 1) * nyp + j) * nzd + k]);
    } else if (t < nBx + nBy + nBz) {
        const int64_t e = t - nBx - nBy;
        const int64_t i = e / (nyd * nzp), j = (e / nzp) % nyd, k = e % nzp;
        if (i < 1 || i >= nxd - 1 || j < 1 || j >= nyd - 1) return;
        Bz[e] += -dtx * (Ey[(i * nyd + j) * nzp + k] - Ey[((i - 1) * nyd + j) * nzp + k]) +
                 dty * (Ex[(i * nyp + j) * nzp + k] - Ex[(i * nyp + j - 1) * nzp + k]);
    }
}

int op_fail(int code, const std::string &m)
{
    set_error(m);
    return code;
}

#define OP_CHECK()                                                              \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) return op_fail(SMI_ERR_HIP, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

int smi_interpolate_3d(int order, const double *Ex, const double *Ey, const double *Ez, const double *Bxm,
                       const double *Bym, const double *Bzm, const int np[3], const int dom_begin[3],
                       const double cell_length[3], int64_t n, const double *x, const double *y, const double *z,
                       double *Epart, double *Bpart, int32_t *iold, double *deltaold, void *stream)
{
    if (order != 2 && order != 4) return op_fail(SMI_ERR_INVALID, "order must be 2 or 4");
    if (n <= 0) return SMI_OK;
    Ints3 NP{{np[0], np[1], np[2]}}, D{{dom_begin[0], dom_begin[1], dom_begin[2]}};
    Dbl3 di{{1.0 / cell_length[0], 1.0 / cell_length[1], 1.0 / cell_length[2]}};
    hipStream_t st = (hipStream_t)stream;
    if (order == 2)
        hipLaunchKernelGGL(k_interp3d<2>, dim3(nblk(n, 256)), dim3(256), 0, st, Ex, Ey, Ez, Bxm, Bym, Bzm, NP, D, di,
                           n, x, y, z, Epart, Bpart, iold, deltaold);
    else
        hipLaunchKernelGGL(k_interp3d<4>, dim3(nblk(n, 256)), dim3(256), 0, st, Ex, Ey, Ez, Bxm, Bym, Bzm, NP, D, di,
                           n, x, y, z, Epart, Bpart, iold, deltaold);
    OP_CHECK();
    return SMI_OK;
}

int smi_interpolate_2d(int order, const double *Ex, const double *Ey, const double *Ez, const double *Bxm,
                       const double *Bym, const double *Bzm, const int np[2], const int dom_begin[2],
                       const double cell_length[2], int64_t n, const double *x, const double *y, double *Epart,
                       double *Bpart, int32_t *iold, double *deltaold, void *stream)
{
    if (order != 2) return op_fail(SMI_ERR_INVALID, "2D interpolation order must be 2");
    if (n <= 0) return SMI_OK;
    Ints3 NP{{np[0], np[1], 1}}, D{{dom_begin[0], dom_begin[1], 0}};
    Dbl3 di{{1.0 / cell_length[0], 1.0 / cell_length[1], 1.0}};
    hipLaunchKernelGGL(k_interp2d, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, Ex, Ey, Ez, Bxm, Bym, Bzm,
                       NP, D, di, n, x, y, Epart, Bpart, iold, deltaold);
    OP_CHECK();
    return SMI_OK;
}

int smi_push_boris(int ndim, double charge, double mass, double timestep, int64_t n, double *x, double *y, double *z,
                   double *px, double *py, double *pz, const double *Epart, const double *Bpart, double *invgf,
                   void *stream)
{
    if (!(mass > 0)) return op_fail(SMI_ERR_INVALID, "mass must be > 0");
    if (n <= 0) return SMI_OK;
    const double one_over_mass = 1.0 / mass, dts2 = timestep / 2.;
    const double qmh = charge * one_over_mass * dts2;
    hipLaunchKernelGGL(k_boris, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, ndim, qmh, timestep, n, x, y,
                       z, px, py, pz, Epart, Bpart, invgf);
    OP_CHECK();
    return SMI_OK;
}

int smi_currents_3d(int order, double *Jx, double *Jy, double *Jz, const int np[3], const int dom_begin[3],
                    const double cell_length[3], double timestep, double charge, int64_t n, const double *x,
                    const double *y, const double *z, const double *w, const int32_t *iold, const double *deltaold,
                    int deterministic, void *stream)
{
    if (order != 2 && order != 4) return op_fail(SMI_ERR_INVALID, "order must be 2 or 4");
    if (n <= 0) return SMI_OK;
    Proj3Args a;
    a.J[0] = Jx;
    a.J[1] = Jy;
    a.J[2] = Jz;
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) a.dims[c][d] = np[d] + (c == d ? 1 : 0);
    for (int d = 0; d < 3; ++d) {
        a.dom.v[d] = dom_begin[d];
        a.dinv.v[d] = 1.0 / cell_length[d];
        a.cr_scale.v[d] = cell_length[d] / timestep;
    }
    double cell_volume = 1.0;
    for (int d = 0; d < 3; ++d) cell_volume *= cell_length[d];
    a.cw_scale = (1. / cell_volume) * charge;
    a.n = n;
    a.x[0] = x;
    a.x[1] = y;
    a.x[2] = z;
    a.w = w;
    a.iold = iold;
    a.dold = deltaold;
    hipStream_t st = (hipStream_t)stream;
    if (deterministic) {
        if (order == 2) hipLaunchKernelGGL(k_proj3d_ordered<2>, dim3(1), dim3(64), 0, st, a);
        else hipLaunchKernelGGL(k_proj3d_ordered<4>, dim3(1), dim3(64), 0, st, a);
    } else {
        if (order == 2) hipLaunchKernelGGL(k_proj3d_atomic<2>, dim3(nblk(n, 128)), dim3(128), 0, st, a);
        else hipLaunchKernelGGL(k_proj3d_atomic<4>, dim3(nblk(n, 128)), dim3(128), 0, st, a);
    }
    OP_CHECK();
    return SMI_OK;
}

int smi_currents_2d(int order, double *Jx, double *Jy, double *Jz, const int np[2], const int dom_begin[2],
                    const double cell_length[2], double timestep, double charge, int64_t n, const double *x,
                    const double *y, const double *pz, const double *w, const double *invgf, const int32_t *iold,
                    const double *deltaold, int deterministic, void *stream)
{
    if (order != 2) return op_fail(SMI_ERR_INVALID, "2D projection order must be 2");
    if (n <= 0) return SMI_OK;
    Proj2Args a;
    a.J[0] = Jx;
    a.J[1] = Jy;
    a.J[2] = Jz;
    a.ny[0] = np[1];
    a.ny[1] = np[1] + 1;
    a.ny[2] = np[1];
    a.nx = np[0];
    a.dom = Ints3{{dom_begin[0], dom_begin[1], 0}};
    a.dinv = Dbl3{{1.0 / cell_length[0], 1.0 / cell_length[1], 1.0}};
    a.cr_scale = Dbl3{{cell_length[0] / timestep, cell_length[1] / timestep, 0.}};
    double cell_volume = 1.0;
    cell_volume *= cell_length[0];
    cell_volume *= cell_length[1];
    a.cw_scale = (1. / cell_volume) * charge;
    a.n = n;
    a.x = x;
    a.y = y;
    a.pz = pz;
    a.w = w;
    a.invgf = invgf;
    a.iold = iold;
    a.dold = deltaold;
    hipStream_t st = (hipStream_t)stream;
    if (deterministic) hipLaunchKernelGGL(k_proj2d_ordered, dim3(1), dim3(64), 0, st, a);
    else hipLaunchKernelGGL(k_proj2d_atomic, dim3(nblk(n, 128)), dim3(128), 0, st, a);
    OP_CHECK();
    return SMI_OK;
}

int smi_maxwell_ampere_3d(double *Ex, double *Ey, double *Ez, const double *Bx, const double *By, const double *Bz,
                          const double *Jx, const double *Jy, const double *Jz, const int np[3], double timestep,
                          const double cell_length[3], void *stream)
{
    Ints3 NP{{np[0], np[1], np[2]}};
    const int64_t tot = (int64_t)(np[0] + 1) * np[1] * np[2] + (int64_t)np[0] * (np[1] + 1) * np[2] +
                        (int64_t)np[0] * np[1] * (np[2] + 1);
    hipLaunchKernelGGL(k_ma3d_patch, dim3(nblk(tot, 256)), dim3(256), 0, (hipStream_t)stream, Ex, Ey, Ez, Bx, By, Bz,
                       Jx, Jy, Jz, NP, timestep, timestep / cell_length[0], timestep / cell_length[1],
                       timestep / cell_length[2]);
    OP_CHECK();
    return SMI_OK;
}

int smi_maxwell_faraday_3d(double *Bx, double *By, double *Bz, const double *Ex, const double *Ey, const double *Ez,
                           const int np[3], double timestep, const double cell_length[3], void *stream)
{
    Ints3 NP{{np[0], np[1], np[2]}};
    const int64_t tot = (int64_t)np[0] * (np[1] + 1) * (np[2] + 1) + (int64_t)(np[0] + 1) * np[1] * (np[2] + 1) +
                        (int64_t)(np[0] + 1) * (np[1] + 1) * np[2];
    hipLaunchKernelGGL(k_mf3d_patch, dim3(nblk(tot, 256)), dim3(256), 0, (hipStream_t)stream, Bx, By, Bz, Ex, Ey, Ez,
                       NP, timestep / cell_length[0], timestep / cell_length[1], timestep / cell_length[2]);
    OP_CHECK();
    return SMI_OK;
}

}  // extern "C"
