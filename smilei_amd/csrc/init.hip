// init.hip -- seeded in-HBM particle initialisation.
//
// Plays the role of ParticleCreator (src/Particles/ParticleCreator.cpp) for
// the synthetic workloads: `ppc` particles per active cell (or a per-cell
// particle list built on the host from density / particles-per-cell
// profiles), positions uniform inside the cell, Gaussian momenta, weight
// density*V_cell/ppc (ParticleCreator.cpp:259, createWeight :933-939).  Random numbers come from a counter-based hash
// (SplitMix64 finaliser of seed, species, particle and draw index), so the
// result does not depend on the launch shape.
#include "kernels.h"
#include "pic_device.h"

namespace smi {

namespace {

using dev::cluster_key3;

__device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform double in (0, 1).
__device__ inline double u01(uint64_t seed, uint64_t i, uint32_t k)
{
    const uint64_t h = mix64(seed ^ mix64(i * 16ull + k));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Momentum of particle i: with a quantile table uq[0..nq) of |u| (a
// Maxwell-Juttner distribution tabulated on the host) an isotropic vector of
// that length, otherwise Gaussian components of std sigma.  A drift (u_d =
// g V along x, g = sqrt(1 + u_d^2)) Lorentz-boosts that rest-frame momentum
// as ParticleCreator's boostParticles does (ParticleCreator.cpp:755-801):
// first Zenitani's volume-transform correction (a particle whose rest-frame
// velocity v_x points against the boost is reflected along it with
// probability -V v_x, so that the boosted sample follows the drifting
// distribution and not the rest-frame one), then u_x' = g u_x + gamma u_d
// (a cold particle gets exactly u_d).
__device__ inline void draw_momentum(uint64_t seed, int64_t i, double sigma, const double *uq, int nq, double drift,
                                     double *u)
{
    if (uq) {
        const double r = u01(seed, i, 3) * (nq - 1);
        const int k = min((int)r, nq - 2);
        const double mag = uq[k] + (r - k) * (uq[k + 1] - uq[k]);
        const double ct = 2.0 * u01(seed, i, 4) - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct));
        const double ph = 6.283185307179586 * u01(seed, i, 5);
        u[0] = mag * st * cos(ph);
        u[1] = mag * st * sin(ph);
        u[2] = mag * ct;
    } else {
        for (int d = 0; d < 3; ++d) {
            const double r1 = u01(seed, i, 3 + 2 * d), r2 = u01(seed, i, 4 + 2 * d);
            u[d] = sigma * sqrt(-2.0 * log(r1)) * cos(6.283185307179586 * r2);
        }
    }
    if (drift != 0.0) {
        const double gb = sqrt(1.0 + drift * drift), V = drift / gb;
        const double gam = sqrt(1.0 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (-V * u[0] / gam > u01(seed, i, 9)) u[0] = -u[0];
        u[0] = fma(gb, u[0], gam * drift);
    }
}

// Particle i lies in cell cells[i / ppc] (all cells in order when cells is
// null) and has weight weights[i] (or `weight` when weights is null).
__global__ void k_init(Geom g, PartPtrs A, MixPtrs M, uint32_t *__restrict__ keys, int64_t n, int ppc,
                       const int32_t *__restrict__ cells, const double *__restrict__ weights, double weight,
                       double sigma, const double *__restrict__ uq, int nq, double drift, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (weights) weight = weights[i];
    const int64_t c = cells ? (int64_t)cells[i / ppc] : i / ppc;
    const int ck = (int)(c % g.N[2]);
    const int cj = (int)((c / g.N[2]) % g.N[1]);
    const int ci = (int)(c / ((int64_t)g.N[2] * g.N[1]));
    const int cc[3] = {ci, cj, ck};
    double x[3], f[3];
    for (int d = 0; d < 3; ++d) {
        f[d] = d < g.nd ? u01(seed, i, d) : 0.;
        double v = ((double)cc[d] + f[d]) * g.dx[d];
        if (v >= g.xmax[d]) v = g.xmax[d] - 0.5 * g.dx[d] * 1e-9;  // guard rounding at the box end
        x[d] = v;
    }
    double u[3];
    draw_momentum(seed, i, sigma, uq, nq, drift, u);
    if (M.ox) {
        // split-precision form: exact nearest node, FP32 offset of the same
        // draw (f in the upper half of the cell belongs to the next node)
        int nd[3];
        float fo[3];
        for (int d = 0; d < 3; ++d) {
            const int up = f[d] >= 0.5 ? 1 : 0;
            nd[d] = wrapi(cc[d] + up, g.N[d]);
            fo[d] = (float)(f[d] - (double)up);
            if (fo[d] >= 0.5f) fo[d] = 0.49999997f;  // FP32 rounding of f just below 1/2
        }
        keys[i] = (13u << 27) | cluster_key3(g, nd);
        M.ox[i] = fo[0];
        M.oy[i] = fo[1];
        M.oz[i] = fo[2];
        M.ux[i] = (float)u[0];
        M.uy[i] = (float)u[1];
        M.uz[i] = (float)u[2];
        M.w[i] = (float)weight;
        return;
    }
    A.x[i] = x[0];
    A.y[i] = x[1];
    A.z[i] = g.nd == 3 ? x[2] : 0.;
    A.px[i] = u[0];
    A.py[i] = u[1];
    A.pz[i] = u[2];
    A.w[i] = weight;
}

// Particles at the positions (and, scaled, the weights) of an existing
// species, fresh Gaussian momenta: Smilei's position_initialization = "<other
// species>" (ParticleCreator.cpp:277-284).  Buffer order is kept; keys are
// the source's cells with direction 13.
__global__ void k_init_like(PartPtrs S, MixPtrs SM, const uint32_t *__restrict__ skeys, PartPtrs D, MixPtrs DM,
                            uint32_t *__restrict__ dkeys, int64_t n, int nd, double wscale, double sigma,
                            const double *__restrict__ uq, int nq, double drift, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[3];
    draw_momentum(seed, i, sigma, uq, nq, drift, u);
    dkeys[i] = (13u << 27) | (skeys[i] & ((1u << 27) - 1));
    if (DM.ox) {
        DM.ox[i] = SM.ox[i];
        DM.oy[i] = SM.oy[i];
        DM.oz[i] = SM.oz[i];
        DM.ux[i] = (float)u[0];
        DM.uy[i] = (float)u[1];
        DM.uz[i] = (float)u[2];
        DM.w[i] = (float)((double)SM.w[i] * wscale);
        return;
    }
    D.x[i] = S.x[i];
    D.y[i] = S.y[i];
    D.z[i] = nd == 3 ? S.z[i] : 0.;
    D.px[i] = u[0];
    D.py[i] = u[1];
    D.pz[i] = u[2];
    D.w[i] = S.w[i] * wscale;
}

}  // namespace

void launch_init_like(const Geom &g, const PartPtrs &S, const MixPtrs &SM, const uint32_t *skeys, const PartPtrs &D,
                      const MixPtrs &DM, uint32_t *dkeys, int64_t n, double wscale, double sigma, const double *uq,
                      int nq, double drift, uint64_t seed, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_init_like, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, S, SM, skeys, D, DM, dkeys, n,
                       g.nd, wscale, sigma, uq, nq, drift, seed);
    SMI_LAUNCH_CHECK();
}

void launch_init(const Geom &g, const PartPtrs &A, const MixPtrs &M, uint32_t *keys, int64_t n, int ppc,
                 const int32_t *cells, const double *weights, double weight, double sigma, const double *uq, int nq,
                 double drift, uint64_t seed, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, A, M, keys, n, ppc, cells,
                       weights, weight, sigma, uq, nq, drift, seed);
    SMI_LAUNCH_CHECK();
}

}  // namespace smi
