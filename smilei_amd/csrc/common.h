// common.h -- shared host/device definitions of the MI355X PIC library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace smi {

// Error raised by host code; converted to an SMI_ERR_* code at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SMI_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::smi::Error(-2, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                       __FILE__ + ":" + std::to_string(__LINE__));           \
    } while (0)

#define SMI_LAUNCH_CHECK() SMI_HIP(hipGetLastError())

// Geometry of the simulation as seen by every kernel (passed by value).
struct Geom {
    int nd;           // 2 or 3
    int order;        // 2 or 4
    int os;           // oversize (ghost cells of the per-patch layout) = order
    int N[3];         // global cells per dimension (1 beyond nd)
    int n[3];         // cells per patch
    int npatch[3];    // patches per dimension
    int ncl[3];       // 4-cell clusters per patch per dimension
    int cpp;          // cells per patch
    int clpp;         // clusters per patch
    int cells_per_cl; // 64 in 3D, 16 in 2D
    int npatches;     // total patches
    int64_t ncells;   // total cells
    double dx[3], dinv[3], dt;
    double dxdt[3];   // dx/dt (the mixed deposit's current factor)
    double plen[3];   // patch length n*dx, as Patch::initStep3 computes it
    double xmax[3];   // global length dx*N, as Patch::prepareParticles does
};

// Physical constants of one species for the particle kernels.
struct SpeciesK {
    double charge, mass;
    double qmh;        // charge * (1/mass) * (dt/2)   (PusherBoris.cpp:86)
    double cw_scale;   // inv_cell_volume * charge     (Projector3D2Order.cpp:64)
};

__host__ __device__ inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

}  // namespace smi

// ----------------------------------------------------------------------------
// B-spline weights with the reference's association order, FP64.
// The library is compiled with -ffp-contract=off so that the products and
// sums round exactly as in the C++ operators they restate; kernels that want
// fused multiply-adds (mixed precision) spell them with fma()/fmaf().
// ----------------------------------------------------------------------------

namespace smi {

// Quadratic spline at offsets -1,0,1 (Interpolator3D2Order.h:108-112).
__host__ __device__ inline void spline2(double d, double *w)
{
    const double d2 = d * d;
    w[0] = 0.5 * (d2 - d + 0.25);
    w[1] = 0.75 - d2;
    w[2] = 0.5 * (d2 + d + 0.25);
}

// Quartic spline at offsets -2..2 (Interpolator3D4Order.h:62-67, constants
// Interpolator3D4Order.cpp:24-34).
__host__ __device__ inline void spline4(double d, double *w)
{
    const double c384 = 1.0 / 384.0, c48 = 1.0 / 48.0, c16 = 1.0 / 16.0, c12 = 1.0 / 12.0,
                 c24 = 1.0 / 24.0, c19_96 = 19.0 / 96.0, c11_24 = 11.0 / 24.0, c4 = 1.0 / 4.0,
                 c6 = 1.0 / 6.0, c115_192 = 115.0 / 192.0, c5_8 = 5.0 / 8.0;
    const double d2 = d * d, d3 = d2 * d, d4 = d3 * d;
    w[0] = c384 - c48 * d + c16 * d2 - c12 * d3 + c24 * d4;
    w[1] = c19_96 - c11_24 * d + c4 * d2 + c6 * d3 - c6 * d4;
    w[2] = c115_192 - c5_8 * d2 + c4 * d4;
    w[3] = c19_96 + c11_24 * d + c4 * d2 - c6 * d3 - c6 * d4;
    w[4] = c384 + c48 * d + c16 * d2 + c12 * d3 + c24 * d4;
}

template <int ORDER>
__host__ __device__ inline void spline(double d, double *w)
{
    if constexpr (ORDER == 2) spline2(d, w);
    else spline4(d, w);
}

}  // namespace smi
