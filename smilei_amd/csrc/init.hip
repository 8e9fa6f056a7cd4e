// init.hip -- seeded in-HBM particle initialisation.
//
// Plays the role of ParticleCreator (src/Particles/ParticleCreator.cpp) for
// the synthetic workloads: `ppc` particles per active cell, positions
// uniform inside the cell, Gaussian momenta, weight density*V_cell/ppc
// (ParticleCreator.cpp:259).  Random numbers come from a counter-based hash
// (SplitMix64 finaliser of seed, species, particle and draw index), so the
// result does not depend on the launch shape.
#include "kernels.h"

namespace smi {

namespace {

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

__global__ void k_init(Geom g, PartPtrs A, int64_t n, int ppc, const int32_t *__restrict__ cells, double weight,
                       double sigma, double drift, uint64_t seed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = cells ? (int64_t)cells[i / ppc] : i / ppc;
    const int ck = (int)(c % g.N[2]);
    const int cj = (int)((c / g.N[2]) % g.N[1]);
    const int ci = (int)(c / ((int64_t)g.N[2] * g.N[1]));
    const int cc[3] = {ci, cj, ck};
    double x[3];
    for (int d = 0; d < 3; ++d) {
        double v = ((double)cc[d] + u01(seed, i, d)) * g.dx[d];
        if (v >= g.xmax[d]) v = g.xmax[d] - 0.5 * g.dx[d] * 1e-9;  // guard rounding at the box end
        x[d] = v;
    }
    double u[3];
    for (int d = 0; d < 3; ++d) {
        const double r1 = u01(seed, i, 3 + 2 * d), r2 = u01(seed, i, 4 + 2 * d);
        u[d] = sigma * sqrt(-2.0 * log(r1)) * cos(6.283185307179586 * r2);
    }
    A.x[i] = x[0];
    A.y[i] = x[1];
    A.z[i] = g.nd == 3 ? x[2] : 0.;
    A.px[i] = u[0] + drift;
    A.py[i] = u[1];
    A.pz[i] = u[2];
    A.w[i] = weight;
}

}  // namespace

void launch_init(const Geom &g, const PartPtrs &A, int64_t n, int ppc, const int32_t *cells, double weight,
                 double sigma, double drift, uint64_t seed, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, A, n, ppc, cells, weight, sigma,
                       drift, seed);
    SMI_LAUNCH_CHECK();
}

}  // namespace smi
