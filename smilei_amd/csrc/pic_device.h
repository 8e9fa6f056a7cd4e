// pic_device.h -- device arithmetic shared by the batched kernels and the
// per-operator entry points: Boris push, B-spline gathers and the Esirkepov
// current contributions, each in the association order of the reference
// operator it restates (the library is built with -ffp-contract=off).
#pragma once

#include "common.h"

namespace smi {
namespace dev {

constexpr double kThird = 1. / 3.;

// Tile geometry.  A cluster spans cells [c0, c0+4) along each axis; the
// nearest primal node of a particle is in [c0, c0+4], the dual one in
// [c0+1, c0+5].  Gather stencil +-H, deposit stencil +-(H+1).
template <int ORDER> struct Tile {
    static constexpr int H = ORDER / 2;
    static constexpr int GO = H;             // gather tile starts at c0 - GO
    static constexpr int GT = 4 + 2 * H + 2; // covers [c0-H, c0+5+H]
    static constexpr int JO = H + 1;         // J tile starts at c0 - JO
    static constexpr int JT = 4 + 2 * H + 3; // covers [c0-H-1, c0+4+H+1]
    static constexpr int S = 2 * H + 3;      // Esirkepov stencil length
};

// Pointers into device memory, in the global address space.  The particle
// arrays reach the kernels through descriptors in memory, which leaves the
// compiler generic (flat) pointers: their loads and stores need a 64-bit
// address in VGPRs per access and count against both the memory and the
// LDS wait counters.  Cast to the global address space, they become global
// loads and stores with a scalar base and a 32-bit offset.
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ inline gptr<T> glb(T *p)
{
    return (gptr<T>)p;
}
struct GMix {
    gptr<float> ox, oy, oz, ux, uy, uz, w;
};
__device__ inline GMix glb(const MixPtrs &M)
{
    return GMix{glb(M.ox), glb(M.oy), glb(M.oz), glb(M.ux), glb(M.uy), glb(M.uz), glb(M.w)};
}
struct GPart {
    gptr<double> x, y, z, px, py, pz, w;
};
__device__ inline GPart glb(const PartPtrs &M)
{
    return GPart{glb(M.x), glb(M.y), glb(M.z), glb(M.px), glb(M.py), glb(M.pz), glb(M.w)};
}

// XCD-aware order of the one-workgroup-per-cluster grid.  Workgroups are
// dispatched round-robin over the 8 XCDs of the chip (workgroup b to XCD
// b % 8), each XCD with its own L2; this mapping gives XCD x the contiguous
// run of clusters [x*n/8, (x+1)*n/8) in order, so that neighbouring clusters
// (which share field-tile halos and J lines) meet in the same L2.  A
// bijection of [0, n) for any n.
constexpr int kXcds = 8;
__device__ inline int xcd_cluster(int b, int n)
{
    const int per = n / kXcds, rem = n % kXcds;
    const int x = b % kXcds, k = b / kXcds;
    return x * per + (x < rem ? x : rem) + k;
}

struct Cluster {
    int p;          // patch id
    int pc[3];      // patch coordinates
    int G0[3];      // first global cell of the cluster
    int64_t gbase;  // global cell key of the cluster's first cell
};

__device__ inline Cluster decode_cluster(const Geom &g, int b)
{
    Cluster c;
    c.p = b / g.clpp;
    const int cl = b % g.clpp;
    int cc[3];
    cc[0] = cl / (g.ncl[1] * g.ncl[2]);
    cc[1] = (cl / g.ncl[2]) % g.ncl[1];
    cc[2] = cl % g.ncl[2];
    c.pc[0] = c.p / (g.npatch[1] * g.npatch[2]);
    c.pc[1] = (c.p / g.npatch[2]) % g.npatch[1];
    c.pc[2] = c.p % g.npatch[2];
    for (int d = 0; d < 3; ++d) c.G0[d] = c.pc[d] * g.n[d] + 4 * cc[d];
    c.gbase = (int64_t)c.p * g.cpp + (int64_t)cl * g.cells_per_cl;
    return c;
}

// Global key (patch * cells-per-patch + cluster * 64 + 16 i + 4 j + k, the
// layout of the cell keys) of global grid point c[3] (already wrapped).
__device__ inline uint32_t cluster_key3(const Geom &g, const int *c)
{
    int pc[3], l[3];
    for (int d = 0; d < 3; ++d) {
        pc[d] = c[d] / g.n[d];
        l[d] = c[d] - pc[d] * g.n[d];
    }
    const int p = (pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2];
    const int cl = ((l[0] >> 2) * g.ncl[1] + (l[1] >> 2)) * g.ncl[2] + (l[2] >> 2);
    return (uint32_t)((int64_t)p * g.cpp + cl * 64 + 16 * (l[0] & 3) + 4 * (l[1] & 3) + (l[2] & 3));
}

// Mixed form: position xn = x/dx (grid units) -> nearest grid node, wrapped
// into [0, N), and the FP32 offset from it in [-1/2, 1/2).
__device__ inline void split_node(double xn, int N, int &node, float &off)
{
    const double fl = floor(xn + 0.5);
    float f = (float)(xn - fl);
    int c = (int)fl;
    if (f >= 0.5f) {  // rounded up to 1/2 in FP32
        f -= 1.f;
        c += 1;
    }
    if (f < -0.5f) f = -0.5f;
    c %= N;
    node = c < 0 ? c + N : c;
    off = f;
}

// In-patch cell key (Cluster3D::CellKey, nvidiaParticles.cu:639-660) of a
// position inside patch `pc`; clamped against rounding at the upper face.
template <int ND>
__device__ inline int32_t cell_key(const Geom &g, const int *pc, const double *x)
{
    int ic[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        int c = (int)floor(x[d] * (1.0 / g.dx[d])) - pc[d] * g.n[d];
        c = c < 0 ? 0 : (c > g.n[d] - 1 ? g.n[d] - 1 : c);
        ic[d] = c;
    }
    if constexpr (ND == 3) {
        const int cl = ((ic[0] >> 2) * g.ncl[1] + (ic[1] >> 2)) * g.ncl[2] + (ic[2] >> 2);
        return cl * 64 + 16 * (ic[0] & 3) + 4 * (ic[1] & 3) + (ic[2] & 3);
    } else {
        const int cl = (ic[0] >> 2) * g.ncl[1] + (ic[1] >> 2);
        return cl * 16 + 4 * (ic[0] & 3) + (ic[1] & 3);
    }
}

// Boris push (PusherBoris.cpp:80-127), FP64, reference association order.
__device__ inline void boris(double qmh, double dt, const double *E, const double *B, double *u,
                             double *dxp, double &invgf)
{
    double ax = qmh * E[0], ay = qmh * E[1], az = qmh * E[2];
    const double ux = u[0] + ax, uy = u[1] + ay, uz = u[2] + az;
    const double f = qmh / sqrt(1.0 + ux * ux + uy * uy + uz * uz);
    const double tx = f * B[0], ty = f * B[1], tz = f * B[2];
    const double inv_det = 1.0 / (1.0 + tx * tx + ty * ty + tz * tz);
    ax += ((1.0 + tx * tx - ty * ty - tz * tz) * ux + 2.0 * (tx * ty + tz) * uy + 2.0 * (tz * tx - ty) * uz) *
          inv_det;
    ay += (2.0 * (tx * ty - tz) * ux + (1.0 - tx * tx + ty * ty - tz * tz) * uy + 2.0 * (ty * tz + tx) * uz) *
          inv_det;
    az += (2.0 * (tz * tx + ty) * ux + 2.0 * (ty * tz - tx) * uy + (1.0 - tx * tx - ty * ty + tz * tz) * uz) *
          inv_det;
    double gi = 1. / sqrt(1.0 + ax * ax + ay * ay + az * az);
    invgf = gi;
    u[0] = ax;
    u[1] = ay;
    u[2] = az;
    gi *= dt;
    dxp[0] = ax * gi;
    dxp[1] = ay * gi;
    dxp[2] = az * gi;
}

// Gather from an LDS tile of width T: sum_{a,b,c} ((cx*cy)*cz)*f, a outer.
template <int H, int T, int SX = T * T>
__device__ inline double gather3(const double *cx, const double *cy, const double *cz, const double *t,
                                 int i, int j, int k)
{
    double r = 0.;
#pragma unroll
    for (int a = -H; a <= H; ++a)
#pragma unroll
        for (int b = -H; b <= H; ++b)
#pragma unroll
            for (int c = -H; c <= H; ++c)
                r += cx[a + H] * cy[b + H] * cz[c + H] * t[(i + a) * SX + (j + b) * T + (k + c)];
    return r;
}

template <int H, int T>
__device__ inline double gather2(const double *cx, const double *cy, const double *t, int i, int j)
{
    double r = 0.;
#pragma unroll
    for (int a = -H; a <= H; ++a)
#pragma unroll
        for (int b = -H; b <= H; ++b) r += cx[a + H] * cy[b + H] * t[(i + a) * T + (j + b)];
    return r;
}

// Same gathers on a strided array (the reference's per-patch Field3D/2D).
template <int H>
__device__ inline double gather3s(const double *cx, const double *cy, const double *cz, const double *f, int i,
                                  int j, int k, int ny, int nz)
{
    double r = 0.;
#pragma unroll
    for (int a = -H; a <= H; ++a)
#pragma unroll
        for (int b = -H; b <= H; ++b)
#pragma unroll
            for (int c = -H; c <= H; ++c)
                r += cx[a + H] * cy[b + H] * cz[c + H] * f[((int64_t)(i + a) * ny + (j + b)) * nz + (k + c)];
    return r;
}

template <int H>
__device__ inline double gather2s(const double *cx, const double *cy, const double *f, int i, int j, int ny)
{
    double r = 0.;
#pragma unroll
    for (int a = -H; a <= H; ++a)
#pragma unroll
        for (int b = -H; b <= H; ++b) r += cx[a + H] * cy[b + H] * f[(int64_t)(i + a) * ny + (j + b)];
    return r;
}

// Esirkepov shapes.  S0 from the old offset, S1 from the new position,
// shifted by ip - ipo; DS = S1 - S0 (Projector3D2Order.cpp:92-148).
template <int ORDER>
__device__ inline void esirkepov_shapes(double dold, double pn_new, int ipo, double *S0, double *S1,
                                        double *DS)
{
    constexpr int S = Tile<ORDER>::S;
    S0[0] = 0.;
    spline<ORDER>(dold, &S0[1]);
    S0[S - 1] = 0.;
#pragma unroll
    for (int i = 0; i < S; ++i) S1[i] = 0.;
    const int ip = (int)round(pn_new);
    const double d = pn_new - (double)ip;
    double w[5];
    spline<ORDER>(d, w);
    const int sh = ip - ipo;  // -1, 0 or 1
#pragma unroll
    for (int i = 0; i < ORDER + 1; ++i) {
        // S1[sh + 1 + i] = w[i] with compile-time indices only
#pragma unroll
        for (int q = -1; q <= 1; ++q)
            if (q == sh) S1[q + 1 + i] = w[i];
    }
#pragma unroll
    for (int i = 0; i < S; ++i) DS[i] = S1[i] - S0[i];
}

// Transverse Esirkepov weight W(b,c) for the current along axis a, with
// (b,c) the two other axes in cyclic order (Jx:(y,z), Jy:(z,x), Jz:(x,y)),
// Projector3D2Order.cpp:176-330.  With S0 zero at both stencil ends this
// single expression rounds exactly like the reference's four special-cased
// blocks.
__device__ inline double weight3(const double *S0b, const double *DSb, const double *S0c,
                                 const double *DSc, int jb, int jc)
{
    return S0b[jb] * S0c[jc] + 0.5 * DSb[jb] * S0c[jc] + 0.5 * DSc[jc] * S0b[jb] +
           kThird * DSb[jb] * DSc[jc];
}

// Walk the 3D Esirkepov contributions of one particle and call
// add(component, i, j, k, value) with (i,j,k) offsets from (ipo, jpo, kpo)
// minus (H+1).  Order 2: tmp = cr*W, acc -= DS*tmp; order 4:
// acc -= (cr*DS)*W (Projector3D4Order.cpp:190-232).
template <int ORDER, class Add>
__device__ inline void esirkepov3d(const double (&S0)[3][7], const double (&DS)[3][7], const double *cr,
                                   Add add)
{
    constexpr int S = Tile<ORDER>::S;
    // Jx: prefix along x over rows i = 1..S-1, transverse (y=j, z=k)
    for (int j = 0; j < S; ++j)
        for (int k = 0; k < S; ++k) {
            const double W = weight3(S0[1], DS[1], S0[2], DS[2], j, k);
            double acc = 0.;
            if constexpr (ORDER == 2) {
                const double t = cr[0] * W;
                for (int i = 1; i < S; ++i) {
                    acc -= DS[0][i - 1] * t;
                    add(0, i, j, k, acc);
                }
            } else {
                for (int i = 1; i < S; ++i) {
                    acc -= cr[0] * DS[0][i - 1] * W;
                    add(0, i, j, k, acc);
                }
            }
        }
    // Jy: prefix along y, transverse (z=k, x=i)
    for (int i = 0; i < S; ++i)
        for (int k = 0; k < S; ++k) {
            const double W = weight3(S0[2], DS[2], S0[0], DS[0], k, i);
            double acc = 0.;
            if constexpr (ORDER == 2) {
                const double t = cr[1] * W;
                for (int j = 1; j < S; ++j) {
                    acc -= DS[1][j - 1] * t;
                    add(1, i, j, k, acc);
                }
            } else {
                for (int j = 1; j < S; ++j) {
                    acc -= cr[1] * DS[1][j - 1] * W;
                    add(1, i, j, k, acc);
                }
            }
        }
    // Jz: prefix along z, transverse (x=i, y=j)
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            const double W = weight3(S0[0], DS[0], S0[1], DS[1], i, j);
            double acc = 0.;
            if constexpr (ORDER == 2) {
                const double t = cr[2] * W;
                for (int k = 1; k < S; ++k) {
                    acc -= DS[2][k - 1] * t;
                    add(2, i, j, k, acc);
                }
            } else {
                for (int k = 1; k < S; ++k) {
                    acc -= cr[2] * DS[2][k - 1] * W;
                    add(2, i, j, k, acc);
                }
            }
        }
}

// The part of a particle's Esirkepov stencil that can be nonzero: along
// each axis S0 (old position) and S1 (new position, shifted by the cell
// change sh in {-1,0,1}) both lie inside a window of ORDER+2 consecutive
// entries of the reference's ORDER+3-entry stencil, starting at w0 = 0 when
// sh < 0 and at w0 = 1 otherwise.  Outside the window S0 = DS = 0 exactly,
// so the skipped columns contribute exact zeros, and the prefix sums along
// the current's axis start from an exact zero: the values added below are
// bit-identical to the reference's.  The window arrays are filled with
// compile-time indices only (selects, no dynamic register indexing).
template <int ORDER>
__device__ inline void esirkepov_window(double dold, double pn_new, int ipo, double *S0w, double *DSw, int &w0,
                                        int *shift = nullptr)
{
    constexpr int W = ORDER + 2;
    double s[ORDER + 1], nw[ORDER + 1];
    spline<ORDER>(dold, s);
    const int ip = (int)round(pn_new);
    spline<ORDER>(pn_new - (double)ip, nw);
    const int sh = ip - ipo;
    w0 = sh < 0 ? 0 : 1;
    if (shift) *shift = sh;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const double s0_w0 = m == 0 ? 0. : s[m - 1];     // window starting at stencil 0
        const double s0_w1 = m <= ORDER ? s[m] : 0.;     // window starting at stencil 1
        const double s1_le = m <= ORDER ? nw[m] : 0.;    // sh <= 0
        const double s1_gt = m >= 1 ? nw[m - 1] : 0.;    // sh == +1
        const double a = w0 == 0 ? s0_w0 : s0_w1;
        const double b = sh <= 0 ? s1_le : s1_gt;
        S0w[m] = a;
        DSw[m] = b - a;
    }
}

// Windowed walk of the 3D Esirkepov contributions (same values as
// esirkepov3d for every point it visits; the points it skips receive exact
// zeros there).  add(component, i, j, k, value) gets reference stencil
// indices (0..ORDER+2) along the three axes.  NT < W narrows the columns
// across the current to the first NT window entries; a particle whose
// nearest node is unchanged on every axis has S0 = DS = 0 past entry ORDER,
// so NT = ORDER + 1 skips exact zeros only.
template <int ORDER, int NT = ORDER + 2, class Add>
__device__ inline void esirkepov3d_win(const double (&S0)[3][ORDER + 2], const double (&DS)[3][ORDER + 2],
                                       const int *w0, const double *cr, Add add)
{
    constexpr int W = ORDER + 2, S = ORDER + 3;
    // Jx: columns (y, z) of the window, prefix along x
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const double Wt = weight3(S0[1], DS[1], S0[2], DS[2], a, b);
            const double t = cr[0] * Wt;
            double acc = 0.;
#pragma unroll
            for (int r = 1; r <= W; ++r) {
                if constexpr (ORDER == 2) acc -= DS[0][r - 1] * t;
                else acc -= cr[0] * DS[0][r - 1] * Wt;
                if (r < W || w0[0] == 0) add(0, w0[0] + r, w0[1] + a, w0[2] + b, acc);
            }
        }
    // Jy: columns (x, z), prefix along y
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const double Wt = weight3(S0[2], DS[2], S0[0], DS[0], b, a);
            const double t = cr[1] * Wt;
            double acc = 0.;
#pragma unroll
            for (int r = 1; r <= W; ++r) {
                if constexpr (ORDER == 2) acc -= DS[1][r - 1] * t;
                else acc -= cr[1] * DS[1][r - 1] * Wt;
                if (r < W || w0[1] == 0) add(1, w0[0] + a, w0[1] + r, w0[2] + b, acc);
            }
        }
    // Jz: columns (x, y), prefix along z
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const double Wt = weight3(S0[0], DS[0], S0[1], DS[1], a, b);
            const double t = cr[2] * Wt;
            double acc = 0.;
#pragma unroll
            for (int r = 1; r <= W; ++r) {
                if constexpr (ORDER == 2) acc -= DS[2][r - 1] * t;
                else acc -= cr[2] * DS[2][r - 1] * Wt;
                if (r < W || w0[2] == 0) add(2, w0[0] + a, w0[1] + b, w0[2] + r, acc);
            }
        }
    (void)S;
}

// 2D3V Esirkepov of Projector2D2Order::currents (Projector2D2Order.cpp:
// 131-188): Jx prefix along x per column j, Jy prefix along y per row i,
// Jz projected with the time-averaged shape.  add(c, i, j, v).
template <class Add>
__device__ inline void esirkepov2d(const double (&S0)[3][7], const double (&S1)[3][7],
                                   const double (&DS)[3][7], double crx, double cry, double crz, Add add)
{
    double tJx[5] = {0., 0., 0., 0., 0.};
    {
        const double h2 = 0.5 * S1[0][0], h3 = S1[0][0];
        add(2, 0, 0, crz * (S1[1][0] * h3));
        double acc = 0.;
        const double ty = S0[0][0] + 0.5 * DS[0][0];
        for (int j = 1; j < 5; ++j) {
            acc -= cry * DS[1][j - 1] * ty;
            add(1, 0, j, acc);
            add(2, 0, j, crz * (S0[1][j] * h2 + S1[1][j] * h3));
        }
    }
    for (int i = 1; i < 5; ++i) {
        tJx[0] -= crx * DS[0][i - 1] * (0.5 * DS[1][0]);
        add(0, i, 0, tJx[0]);
        const double h2 = 0.5 * S1[0][i] + S0[0][i], h3 = 0.5 * S0[0][i] + S1[0][i];
        add(2, i, 0, crz * (S1[1][0] * h3));
        double acc = 0.;
        const double ty = S0[0][i] + 0.5 * DS[0][i];
        for (int j = 1; j < 5; ++j) {
            tJx[j] -= crx * DS[0][i - 1] * (S0[1][j] + 0.5 * DS[1][j]);
            add(0, i, j, tJx[j]);
            acc -= cry * DS[1][j - 1] * ty;
            add(1, i, j, acc);
            add(2, i, j, crz * (S0[1][j] * h2 + S1[1][j] * h3));
        }
    }
}

}  // namespace dev
}  // namespace smi
