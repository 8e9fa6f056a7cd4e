// oracle_ops.cpp -- single-patch PIC operators of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see oracle.h).
// Every function restates the arithmetic of the cited reference operator in
// the same order of floating-point operations, so that a HIP kernel compiled
// without FP contraction can be compared bit for bit.  The code is organised
// differently from the reference (shared spline helpers, one stencil walker
// per component) but each sum and product is evaluated in the reference's
// association order.  Compile with -ffp-contract=off.
#include "oracle.h"

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace ora {

static constexpr double kThird = 1. / 3.;

// Quadratic B-spline weights at offsets -1,0,+1 of a node at distance
// `d` (|d| <= 1/2).  Interpolator3D2Order.h:108-112 / Projector3D2Order.cpp:98-102.
static inline void spline2(double d, double *w)
{
    const double d2 = d * d;
    w[0] = 0.5 * (d2 - d + 0.25);
    w[1] = 0.75 - d2;
    w[2] = 0.5 * (d2 + d + 0.25);
}

// Quartic B-spline weights at offsets -2..+2.
// Interpolator3D4Order.h:62-67 (constants Interpolator3D4Order.cpp:24-34).
static const double c384 = 1.0 / 384.0, c48 = 1.0 / 48.0, c16 = 1.0 / 16.0, c12 = 1.0 / 12.0,
                    c24 = 1.0 / 24.0, c19_96 = 19.0 / 96.0, c11_24 = 11.0 / 24.0, c4 = 1.0 / 4.0,
                    c6 = 1.0 / 6.0, c115_192 = 115.0 / 192.0, c5_8 = 5.0 / 8.0;
static inline void spline4(double d, double *w)
{
    const double d2 = d * d, d3 = d2 * d, d4 = d3 * d;
    w[0] = c384 - c48 * d + c16 * d2 - c12 * d3 + c24 * d4;
    w[1] = c19_96 - c11_24 * d + c4 * d2 + c6 * d3 - c6 * d4;
    w[2] = c115_192 - c5_8 * d2 + c4 * d4;
    w[3] = c19_96 + c11_24 * d + c4 * d2 - c6 * d3 - c6 * d4;
    w[4] = c384 + c48 * d + c16 * d2 + c12 * d3 + c24 * d4;
}

static inline void spline(int order, double d, double *w)
{
    if (order == 2) spline2(d, w);
    else spline4(d, w);
}

// 3D gather of one staggered component: sum_{i,j,k} ((cx*cy)*cz)*f, i outer,
// k inner (Interpolator3D2Order.h:54-84, Interpolator3D4Order.h:44-55).
static inline double gather3(int h, const double *cx, const double *cy, const double *cz,
                             const double *f, int ix, int iy, int iz, int ny, int nz)
{
    double r = 0.;
    for (int a = -h; a <= h; ++a)
        for (int b = -h; b <= h; ++b)
            for (int c = -h; c <= h; ++c)
                r += cx[a + h] * cy[b + h] * cz[c + h] *
                     f[((size_t)(ix + a) * ny + (iy + b)) * nz + (iz + c)];
    return r;
}

static inline double gather2(int h, const double *cx, const double *cy, const double *f,
                             int ix, int iy, int ny)
{
    double r = 0.;
    for (int a = -h; a <= h; ++a)
        for (int b = -h; b <= h; ++b)
            r += cx[a + h] * cy[b + h] * f[(size_t)(ix + a) * ny + (iy + b)];
    return r;
}

}  // namespace ora

using namespace ora;

extern "C" {

// Interpolator3D2Order::fieldsWrapper (Interpolator3D2Order.cpp:163-285) and
// Interpolator3D4Order::fieldsWrapper (Interpolator3D4Order.cpp:159-).
// Writes Epart/Bpart/iold/deltaold with stride n like SmileiMPI's buffers.
void ora_interp3d(int order, const double *Ex, const double *Ey, const double *Ez,
                  const double *Bxm, const double *Bym, const double *Bzm,
                  const int np[3], const int dom_begin[3], const double dinv[3],
                  int64_t n, const double *x, const double *y, const double *z,
                  double *Epart, double *Bpart, int32_t *iold, double *deltaold)
{
    const int h = order / 2;
    const int nxp = np[0], nyp = np[1], nzp = np[2];
    (void)nxp;
    const int nyd = nyp + 1, nzd = nzp + 1;
    for (int64_t ip = 0; ip < n; ++ip) {
        const double pn[3] = {x[ip] * dinv[0], y[ip] * dinv[1], z[ip] * dinv[2]};
        int ipr[3], idu[3];
        double cp[3][5], cd[3][5], dp[3];
        for (int d = 0; d < 3; ++d) {
            ipr[d] = (int)std::round(pn[d]);
            dp[d] = pn[d] - (double)ipr[d];
            spline(order, dp[d], cp[d]);
            idu[d] = (int)std::round(pn[d] + 0.5);
            spline(order, pn[d] - (double)idu[d] + 0.5, cd[d]);
            ipr[d] -= dom_begin[d];
            idu[d] -= dom_begin[d];
        }
        // Ex(d,p,p) Ey(p,d,p) Ez(p,p,d) Bx(p,d,d) By(d,p,d) Bz(d,d,p)
        Epart[0 * n + ip] = gather3(h, cd[0], cp[1], cp[2], Ex, idu[0], ipr[1], ipr[2], nyp, nzp);
        Epart[1 * n + ip] = gather3(h, cp[0], cd[1], cp[2], Ey, ipr[0], idu[1], ipr[2], nyd, nzp);
        Epart[2 * n + ip] = gather3(h, cp[0], cp[1], cd[2], Ez, ipr[0], ipr[1], idu[2], nyp, nzd);
        Bpart[0 * n + ip] = gather3(h, cp[0], cd[1], cd[2], Bxm, ipr[0], idu[1], idu[2], nyd, nzd);
        Bpart[1 * n + ip] = gather3(h, cd[0], cp[1], cd[2], Bym, idu[0], ipr[1], idu[2], nyp, nzd);
        Bpart[2 * n + ip] = gather3(h, cd[0], cd[1], cp[2], Bzm, idu[0], idu[1], ipr[2], nyd, nzp);
        for (int d = 0; d < 3; ++d) {
            iold[d * n + ip] = ipr[d];
            deltaold[d * n + ip] = dp[d];
        }
    }
}

// Interpolator2D2Order::fieldsWrapper (Interpolator2D2Order.cpp:158-).
void ora_interp2d(int order, const double *Ex, const double *Ey, const double *Ez,
                  const double *Bxm, const double *Bym, const double *Bzm,
                  const int np[2], const int dom_begin[2], const double dinv[2],
                  int64_t n, const double *x, const double *y,
                  double *Epart, double *Bpart, int32_t *iold, double *deltaold)
{
    const int h = order / 2;
    const int nyp = np[1], nyd = nyp + 1;
    for (int64_t ip = 0; ip < n; ++ip) {
        const double pn[2] = {x[ip] * dinv[0], y[ip] * dinv[1]};
        int ipr[2], idu[2];
        double cp[2][5], cd[2][5], dp[2];
        for (int d = 0; d < 2; ++d) {
            ipr[d] = (int)std::round(pn[d]);
            dp[d] = pn[d] - (double)ipr[d];
            spline(order, dp[d], cp[d]);
            idu[d] = (int)std::round(pn[d] + 0.5);
            spline(order, pn[d] - (double)idu[d] + 0.5, cd[d]);
            ipr[d] -= dom_begin[d];
            idu[d] -= dom_begin[d];
        }
        // Ex(d,p) Ey(p,d) Ez(p,p) Bx(p,d) By(d,p) Bz(d,d)
        Epart[0 * n + ip] = gather2(h, cd[0], cp[1], Ex, idu[0], ipr[1], nyp);
        Epart[1 * n + ip] = gather2(h, cp[0], cd[1], Ey, ipr[0], idu[1], nyd);
        Epart[2 * n + ip] = gather2(h, cp[0], cp[1], Ez, ipr[0], ipr[1], nyp);
        Bpart[0 * n + ip] = gather2(h, cp[0], cd[1], Bxm, ipr[0], idu[1], nyd);
        Bpart[1 * n + ip] = gather2(h, cd[0], cp[1], Bym, idu[0], ipr[1], nyp);
        Bpart[2 * n + ip] = gather2(h, cd[0], cd[1], Bzm, idu[0], idu[1], nyd);
        for (int d = 0; d < 2; ++d) {
            iold[d * n + ip] = ipr[d];
            deltaold[d * n + ip] = dp[d];
        }
    }
}

// PusherBoris::operator() (PusherBoris.cpp:23-142) with the constants of
// Pusher::Pusher (Pusher.cpp:5-29): one_over_mass = 1/m, dts2 = dt/2.
void ora_push_boris(int ndim, double charge, double mass, double dt, int64_t n,
                    double *x, double *y, double *z,
                    double *px, double *py, double *pz,
                    const double *Epart, const double *Bpart, double *invgf)
{
    const double one_over_mass = 1.0 / mass;
    const double dts2 = dt / 2.;
    for (int64_t i = 0; i < n; ++i) {
        const double qmh = charge * one_over_mass * dts2;
        double ax = qmh * Epart[i], ay = qmh * Epart[n + i], az = qmh * Epart[2 * n + i];
        const double ux = px[i] + ax, uy = py[i] + ay, uz = pz[i] + az;
        const double f = qmh / std::sqrt(1.0 + ux * ux + uy * uy + uz * uz);
        const double tx = f * Bpart[i], ty = f * Bpart[n + i], tz = f * Bpart[2 * n + i];
        const double inv_det = 1.0 / (1.0 + tx * tx + ty * ty + tz * tz);
        ax += ((1.0 + tx * tx - ty * ty - tz * tz) * ux + 2.0 * (tx * ty + tz) * uy +
               2.0 * (tz * tx - ty) * uz) * inv_det;
        ay += (2.0 * (tx * ty - tz) * ux + (1.0 - tx * tx + ty * ty - tz * tz) * uy +
               2.0 * (ty * tz + tx) * uz) * inv_det;
        az += (2.0 * (tz * tx + ty) * ux + 2.0 * (ty * tz - tx) * uy +
               (1.0 - tx * tx - ty * ty + tz * tz) * uz) * inv_det;
        double g = 1. / std::sqrt(1.0 + ax * ax + ay * ay + az * az);
        invgf[i] = g;
        px[i] = ax;
        py[i] = ay;
        pz[i] = az;
        g *= dt;
        x[i] += ax * g;
        if (ndim > 1) {
            y[i] += ay * g;
            if (ndim > 2) z[i] += az * g;
        }
    }
}

// Esirkepov deposit, order 2: Projector3D2Order::currents
// (Projector3D2Order.cpp:55-343).  Each component is a prefix sum along its
// own axis of a 5x5 transverse weight; the reference splits the transverse
// loop into four blocks (first row/column use S1 only) and adds the running
// prefix at rows 1..4.  The blocks below follow the same association order.
static void proj3d_o2(double *Jx, double *Jy, double *Jz, const int np[3],
                      const int dom_begin[3], const double dx[3], double dt,
                      double inv_cell_volume, double charge, double w, const double pos[3],
                      const int io[3], const double dold[3])
{
    const double dinv[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
    const double dov[3] = {dx[0] / dt, dx[1] / dt, dx[2] / dt};
    const double cw = inv_cell_volume * charge * w;
    const double cr[3] = {cw * dov[0], cw * dov[1], cw * dov[2]};
    double S0[3][5], S1[3][5], DS[3][5];
    for (int d = 0; d < 3; ++d) {
        S0[d][0] = 0.;
        spline2(dold[d], &S0[d][1]);
        S0[d][4] = 0.;
        for (int i = 0; i < 5; ++i) S1[d][i] = 0.;
        const double pn = pos[d] * dinv[d];
        const int ip = (int)std::round(pn);
        const int shift = ip - io[d] - dom_begin[d];
        spline2(pn - (double)ip, &S1[d][shift + 1]);
        for (int i = 0; i < 5; ++i) DS[d][i] = S1[d][i] - S0[d][i];
    }
    const int o0 = io[0] - 2, o1 = io[1] - 2, o2 = io[2] - 2;
    const int nyp = np[1], nzp = np[2];

    // ---- Jx (d,p,p): axis x, transverse (y,z); strides yz=nzp*nyp, z=nzp
    {
        const size_t zs = nzp, yzs = (size_t)nzp * nyp;
        const size_t base = (size_t)o0 * yzs + (size_t)o1 * zs + o2;
        // (j,k) = (0,0)
        {
            double acc = 0.;
            const double t = cr[0] * (kThird * S1[1][0] * S1[2][0]);
            for (int i = 1; i < 5; ++i) {
                acc -= DS[0][i - 1] * t;
                Jx[base + i * yzs] += acc;
            }
        }
        // j = 0, k = 1..4
        for (int k = 1; k < 5; ++k) {
            double acc = 0.;
            const double t = cr[0] * (0.5 * S1[1][0] * S0[2][k] + kThird * S1[1][0] * DS[2][k]);
            for (int i = 1; i < 5; ++i) {
                acc -= DS[0][i - 1] * t;
                Jx[base + k + i * yzs] += acc;
            }
        }
        // j = 1..4, k = 0
        for (int j = 1; j < 5; ++j) {
            double acc = 0.;
            const double t = cr[0] * (0.5 * S1[2][0] * S0[1][j] + kThird * DS[1][j] * S1[2][0]);
            for (int i = 1; i < 5; ++i) {
                acc -= DS[0][i - 1] * t;
                Jx[base + j * zs + i * yzs] += acc;
            }
        }
        // j,k = 1..4
        for (int j = 1; j < 5; ++j)
            for (int k = 1; k < 5; ++k) {
                double acc = 0.;
                const double t = cr[0] * (S0[1][j] * S0[2][k] + 0.5 * DS[1][j] * S0[2][k] +
                                          0.5 * DS[2][k] * S0[1][j] + kThird * DS[1][j] * DS[2][k]);
                for (int i = 1; i < 5; ++i) {
                    acc -= DS[0][i - 1] * t;
                    Jx[base + j * zs + k + i * yzs] += acc;
                }
            }
    }
    // ---- Jy (p,d,p): axis y, transverse (x,z); yz stride nzp*(nyp+1)
    {
        const size_t zs = nzp, yzs = (size_t)nzp * (nyp + 1);
        const size_t base = (size_t)o0 * yzs + (size_t)o1 * zs + o2;
        {
            double acc = 0.;
            const double t = cr[1] * (kThird * S1[2][0] * S1[0][0]);
            for (int j = 1; j < 5; ++j) {
                acc -= DS[1][j - 1] * t;
                Jy[base + j * zs] += acc;
            }
        }
        for (int k = 1; k < 5; ++k) {
            double acc = 0.;
            const double t = cr[1] * (0.5 * S1[0][0] * S0[2][k] + kThird * DS[2][k] * S1[0][0]);
            for (int j = 1; j < 5; ++j) {
                acc -= DS[1][j - 1] * t;
                Jy[base + k + j * zs] += acc;
            }
        }
        for (int i = 1; i < 5; ++i) {
            double acc = 0.;
            const double t = cr[1] * (0.5 * S1[2][0] * S0[0][i] + kThird * S1[2][0] * DS[0][i]);
            for (int j = 1; j < 5; ++j) {
                acc -= DS[1][j - 1] * t;
                Jy[base + i * yzs + j * zs] += acc;
            }
        }
        for (int i = 1; i < 5; ++i)
            for (int k = 1; k < 5; ++k) {
                double acc = 0.;
                const double t = cr[1] * (S0[2][k] * S0[0][i] + 0.5 * DS[2][k] * S0[0][i] +
                                          0.5 * DS[0][i] * S0[2][k] + kThird * DS[2][k] * DS[0][i]);
                for (int j = 1; j < 5; ++j) {
                    acc -= DS[1][j - 1] * t;
                    Jy[base + i * yzs + k + j * zs] += acc;
                }
            }
    }
    // ---- Jz (p,p,d): axis z, transverse (x,y); z stride nzp+1
    {
        const size_t zs = nzp + 1, yzs = (size_t)(nzp + 1) * nyp;
        const size_t base = (size_t)o0 * yzs + (size_t)o1 * zs + o2;
        {
            double acc = 0.;
            const double t = cr[2] * (kThird * S1[0][0] * S1[1][0]);
            for (int k = 1; k < 5; ++k) {
                acc -= DS[2][k - 1] * t;
                Jz[base + k] += acc;
            }
        }
        for (int j = 1; j < 5; ++j) {
            double acc = 0.;
            const double t = cr[2] * (0.5 * S1[0][0] * S0[1][j] + kThird * S1[0][0] * DS[1][j]);
            for (int k = 1; k < 5; ++k) {
                acc -= DS[2][k - 1] * t;
                Jz[base + j * zs + k] += acc;
            }
        }
        for (int i = 1; i < 5; ++i) {
            double acc = 0.;
            const double t = cr[2] * (0.5 * S1[1][0] * S0[0][i] + kThird * DS[0][i] * S1[1][0]);
            for (int k = 1; k < 5; ++k) {
                acc -= DS[2][k - 1] * t;
                Jz[base + i * yzs + k] += acc;
            }
        }
        for (int i = 1; i < 5; ++i)
            for (int j = 1; j < 5; ++j) {
                double acc = 0.;
                const double t = cr[2] * (S0[0][i] * S0[1][j] + 0.5 * DS[0][i] * S0[1][j] +
                                          0.5 * DS[1][j] * S0[0][i] + kThird * DS[0][i] * DS[1][j]);
                for (int k = 1; k < 5; ++k) {
                    acc -= DS[2][k - 1] * t;
                    Jz[base + i * yzs + j * zs + k] += acc;
                }
            }
    }
}

// Esirkepov deposit, order 4: Projector3D4Order::currents
// (Projector3D4Order.cpp:49-237): 7-point stencils, running sum
// tmp -= (cr*DS[i-1]) * W over the full 7x7 transverse plane.
static void proj3d_o4(double *Jx, double *Jy, double *Jz, const int np[3],
                      const int dom_begin[3], const double dx[3], double dt,
                      double inv_cell_volume, double charge, double w, const double pos[3],
                      const int io[3], const double dold[3])
{
    const double dinv[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
    const double dov[3] = {dx[0] / dt, dx[1] / dt, dx[2] / dt};
    const double cw = inv_cell_volume * charge * w;
    const double cr[3] = {cw * dov[0], cw * dov[1], cw * dov[2]};
    double S0[3][7], S1[3][7], DS[3][7];
    for (int d = 0; d < 3; ++d) {
        S0[d][0] = 0.;
        spline4(dold[d], &S0[d][1]);
        S0[d][6] = 0.;
        for (int i = 0; i < 7; ++i) S1[d][i] = 0.;
        const double pn = pos[d] * dinv[d];
        const int ip = (int)std::round(pn);
        const int shift = ip - io[d] - dom_begin[d];
        spline4(pn - (double)ip, &S1[d][shift + 1]);
        for (int i = 0; i < 7; ++i) DS[d][i] = S1[d][i] - S0[d][i];
    }
    const int o0 = io[0] - 3, o1 = io[1] - 3, o2 = io[2] - 3;
    const int nyp = np[1], nzp = np[2];
    double t[7][7];
    // Jx (d,p,p)
    for (int j = 0; j < 7; ++j)
        for (int k = 0; k < 7; ++k) t[j][k] = 0.;
    for (int i = 1; i < 7; ++i)
        for (int j = 0; j < 7; ++j)
            for (int k = 0; k < 7; ++k) {
                t[j][k] -= cr[0] * DS[0][i - 1] *
                           (S0[1][j] * S0[2][k] + 0.5 * DS[1][j] * S0[2][k] +
                            0.5 * DS[2][k] * S0[1][j] + kThird * DS[1][j] * DS[2][k]);
                Jx[(size_t)(i + o0) * nzp * nyp + (size_t)(j + o1) * nzp + (k + o2)] += t[j][k];
            }
    // Jy (p,d,p)
    for (int i = 0; i < 7; ++i)
        for (int k = 0; k < 7; ++k) t[i][k] = 0.;
    for (int i = 0; i < 7; ++i)
        for (int j = 1; j < 7; ++j)
            for (int k = 0; k < 7; ++k) {
                t[i][k] -= cr[1] * DS[1][j - 1] *
                           (S0[2][k] * S0[0][i] + 0.5 * DS[2][k] * S0[0][i] +
                            0.5 * DS[0][i] * S0[2][k] + kThird * DS[2][k] * DS[0][i]);
                Jy[(size_t)(i + o0) * nzp * (nyp + 1) + (size_t)(j + o1) * nzp + (k + o2)] += t[i][k];
            }
    // Jz (p,p,d)
    for (int i = 0; i < 7; ++i)
        for (int j = 0; j < 7; ++j) t[i][j] = 0.;
    for (int i = 0; i < 7; ++i)
        for (int j = 0; j < 7; ++j)
            for (int k = 1; k < 7; ++k) {
                t[i][j] -= cr[2] * DS[2][k - 1] *
                           (S0[0][i] * S0[1][j] + 0.5 * DS[0][i] * S0[1][j] +
                            0.5 * DS[1][j] * S0[0][i] + kThird * DS[0][i] * DS[1][j]);
                Jz[(size_t)(i + o0) * (nzp + 1) * nyp + (size_t)(j + o1) * (nzp + 1) + (k + o2)] += t[i][j];
            }
}

void ora_proj3d(int order, double *Jx, double *Jy, double *Jz,
                const int np[3], const int dom_begin[3], const double dx[3],
                double dt, double inv_cell_volume, double charge, int64_t n,
                const double *x, const double *y, const double *z,
                const double *w, const int32_t *iold, const double *deltaold)
{
    for (int64_t ip = 0; ip < n; ++ip) {
        const double pos[3] = {x[ip], y[ip], z[ip]};
        const int io[3] = {iold[ip], iold[n + ip], iold[2 * n + ip]};
        const double dold[3] = {deltaold[ip], deltaold[n + ip], deltaold[2 * n + ip]};
        if (order == 2)
            proj3d_o2(Jx, Jy, Jz, np, dom_begin, dx, dt, inv_cell_volume, charge, w[ip], pos, io, dold);
        else
            proj3d_o4(Jx, Jy, Jz, np, dom_begin, dx, dt, inv_cell_volume, charge, w[ip], pos, io, dold);
    }
}

// Esirkepov deposit in 2D3V, order 2: Projector2D2Order::currents
// (Projector2D2Order.cpp:49-189).  Jz (p,p) is projected with the
// time-averaged shape and u_z/gamma (crz = cw/3 * u_z * invgf).
void ora_proj2d(int order, double *Jx, double *Jy, double *Jz,
                const int np[2], const int dom_begin[2], const double dx[2],
                double dt, double inv_cell_volume, double charge, int64_t n,
                const double *x, const double *y, const double *pz,
                const double *w, const double *invgf,
                const int32_t *iold, const double *deltaold)
{
    (void)order;
    const double dinv[2] = {1.0 / dx[0], 1.0 / dx[1]};
    const double dov[2] = {dx[0] / dt, dx[1] / dt};
    const int nyp = np[1];
    for (int64_t ip = 0; ip < n; ++ip) {
        const double cw = inv_cell_volume * charge * w[ip];
        const double crx = cw * dov[0], cry = cw * dov[1];
        const double crz = cw * kThird * pz[ip] * invgf[ip];
        double S0[2][5], S1[2][5], DS[2][5];
        const double pos[2] = {x[ip], y[ip]};
        for (int d = 0; d < 2; ++d) {
            S0[d][0] = 0.;
            S0[d][4] = 0.;
            spline2(deltaold[d * n + ip], &S0[d][1]);
            for (int i = 0; i < 5; ++i) S1[d][i] = 0.;
            const double pn = pos[d] * dinv[d];
            const int ipn = (int)std::round(pn);
            spline2(pn - (double)ipn, &S1[d][ipn - iold[d * n + ip] - dom_begin[d] + 1]);
            for (int i = 0; i < 5; ++i) DS[d][i] = S1[d][i] - S0[d][i];
        }
        const int o0 = iold[ip] - 2, o1 = iold[n + ip] - 2;
        double tJx[5] = {0., 0., 0., 0., 0.};
        // row i = 0: Jx receives nothing (prefix starts at i=1)
        {
            const size_t row = (size_t)o0 * nyp + o1;
            const double h2 = 0.5 * S1[0][0], h3 = S1[0][0];
            Jz[row] += crz * (S1[1][0] * h3);
            double acc = 0.;
            const double ty = S0[0][0] + 0.5 * DS[0][0];
            for (int j = 1; j < 5; ++j) {
                acc -= cry * DS[1][j - 1] * ty;
                Jy[row + j + o0] += acc;
                Jz[row + j] += crz * (S0[1][j] * h2 + S1[1][j] * h3);
            }
        }
        for (int i = 1; i < 5; ++i) {
            const size_t row = (size_t)(i + o0) * nyp + o1;
            tJx[0] -= crx * DS[0][i - 1] * (0.5 * DS[1][0]);
            Jx[row] += tJx[0];
            const double h2 = 0.5 * S1[0][i] + S0[0][i], h3 = 0.5 * S0[0][i] + S1[0][i];
            Jz[row] += crz * (S1[1][0] * h3);
            double acc = 0.;
            const double ty = S0[0][i] + 0.5 * DS[0][i];
            for (int j = 1; j < 5; ++j) {
                tJx[j] -= crx * DS[0][i - 1] * (S0[1][j] + 0.5 * DS[1][j]);
                Jx[row + j] += tJx[j];
                acc -= cry * DS[1][j - 1] * ty;
                Jy[row + j + (i + o0)] += acc;
                Jz[row + j] += crz * (S0[1][j] * h2 + S1[1][j] * h3);
            }
        }
    }
}

// Charge density of one species (type 0 of Projector3D2Order::basic,
// Projector3D2Order.cpp:526-621; order 4 uses the quartic weights on a
// 7-point frame in the same way).
void ora_rho3d(int order, double *rho, const int np[3], const int dom_begin[3],
               const double dinv[3], double inv_cell_volume, double charge,
               int64_t n, const double *x, const double *y, const double *z,
               const double *w)
{
    const int h = order / 2, L = 2 * h + 3;
    const int nyp = np[1], nzp = np[2];
    for (int64_t ip = 0; ip < n; ++ip) {
        const double cw = inv_cell_volume * charge * w[ip];
        double S[3][7];
        int o[3];
        const double pos[3] = {x[ip], y[ip], z[ip]};
        for (int d = 0; d < 3; ++d) {
            for (int i = 0; i < L; ++i) S[d][i] = 0.;
            const double pn = pos[d] * dinv[d];
            const int ipn = (int)std::round(pn);
            spline(order, pn - (double)ipn, &S[d][1]);
            o[d] = ipn - dom_begin[d] - (h + 1);
        }
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j)
                for (int k = 0; k < L; ++k)
                    rho[((size_t)(o[0] + i) * nyp + (o[1] + j)) * nzp + (o[2] + k)] +=
                        cw * S[0][i] * S[1][j] * S[2][k];
    }
}

void ora_rho2d(int order, double *rho, const int np[2], const int dom_begin[2],
               const double dinv[2], double inv_cell_volume, double charge,
               int64_t n, const double *x, const double *y, const double *w)
{
    const int h = order / 2, L = 2 * h + 3;
    const int nyp = np[1];
    for (int64_t ip = 0; ip < n; ++ip) {
        const double cw = inv_cell_volume * charge * w[ip];
        double S[2][7];
        int o[2];
        const double pos[2] = {x[ip], y[ip]};
        for (int d = 0; d < 2; ++d) {
            for (int i = 0; i < L; ++i) S[d][i] = 0.;
            const double pn = pos[d] * dinv[d];
            const int ipn = (int)std::round(pn);
            spline(order, pn - (double)ipn, &S[d][1]);
            o[d] = ipn - dom_begin[d] - (h + 1);
        }
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < L; ++j)
                rho[(size_t)(o[0] + i) * nyp + (o[1] + j)] += cw * S[0][i] * S[1][j];
    }
}

// MA_Solver3D_norm::operator() (MA_Solver3D_norm.cpp:18-115): all points.
void ora_yee3d_ma(double *Ex, double *Ey, double *Ez,
                  const double *Bx, const double *By, const double *Bz,
                  const double *Jx, const double *Jy, const double *Jz,
                  const int np[3], double dt, const double dx[3])
{
    const double dtx = dt / dx[0], dty = dt / dx[1], dtz = dt / dx[2];
    const size_t nxp = np[0], nyp = np[1], nzp = np[2];
    const size_t nxd = nxp + 1, nyd = nyp + 1, nzd = nzp + 1;
    for (size_t i = 0; i < nxd; ++i)
        for (size_t j = 0; j < nyp; ++j)
            for (size_t k = 0; k < nzp; ++k) {
                const size_t e = (i * nyp + j) * nzp + k;
                Ex[e] += -dt * Jx[e] + dty * (Bz[(i * nyd + j + 1) * nzp + k] - Bz[(i * nyd + j) * nzp + k]) -
                         dtz * (By[(i * nyp + j) * nzd + k + 1] - By[(i * nyp + j) * nzd + k]);
            }
    for (size_t i = 0; i < nxp; ++i)
        for (size_t j = 0; j < nyd; ++j)
            for (size_t k = 0; k < nzp; ++k) {
                const size_t e = (i * nyd + j) * nzp + k;
                Ey[e] += -dt * Jy[e] - dtx * (Bz[((i + 1) * nyd + j) * nzp + k] - Bz[(i * nyd + j) * nzp + k]) +
                         dtz * (Bx[(i * nyd + j) * nzd + k + 1] - Bx[(i * nyd + j) * nzd + k]);
            }
    for (size_t i = 0; i < nxp; ++i)
        for (size_t j = 0; j < nyp; ++j)
            for (size_t k = 0; k < nzd; ++k) {
                const size_t e = (i * nyp + j) * nzd + k;
                Ez[e] += -dt * Jz[e] + dtx * (By[((i + 1) * nyp + j) * nzd + k] - By[(i * nyp + j) * nzd + k]) -
                         dty * (Bx[(i * nyd + j + 1) * nzd + k] - Bx[(i * nyd + j) * nzd + k]);
            }
}

// MF_Solver3D_Yee::operator() (MF_Solver3D_Yee.cpp:17-110): interior of the
// dual axes only; the outer dual planes are filled by the B exchange.
void ora_yee3d_mf(double *Bx, double *By, double *Bz,
                  const double *Ex, const double *Ey, const double *Ez,
                  const int np[3], double dt, const double dx[3])
{
    const double dtx = dt / dx[0], dty = dt / dx[1], dtz = dt / dx[2];
    const size_t nxp = np[0], nyp = np[1], nzp = np[2];
    const size_t nxd = nxp + 1, nyd = nyp + 1, nzd = nzp + 1;
    for (size_t i = 0; i < nxp; ++i)
        for (size_t j = 1; j < nyd - 1; ++j)
            for (size_t k = 1; k < nzd - 1; ++k) {
                const size_t b = (i * nyd + j) * nzd + k;
                Bx[b] += -dty * (Ez[(i * nyp + j) * nzd + k] - Ez[(i * nyp + j - 1) * nzd + k]) +
                         dtz * (Ey[(i * nyd + j) * nzp + k] - Ey[(i * nyd + j) * nzp + k - 1]);
            }
    for (size_t i = 1; i < nxd - 1; ++i)
        for (size_t j = 0; j < nyp; ++j)
            for (size_t k = 1; k < nzd - 1; ++k) {
                const size_t b = (i * nyp + j) * nzd + k;
                By[b] += -dtz * (Ex[(i * nyp + j) * nzp + k] - Ex[(i * nyp + j) * nzp + k - 1]) +
                         dtx * (Ez[(i * nyp + j) * nzd + k] - Ez[((i - 1) * nyp + j) * nzd + k]);
            }
    for (size_t i = 1; i < nxd - 1; ++i)
        for (size_t j = 1; j < nyd - 1; ++j)
            for (size_t k = 0; k < nzp; ++k) {
                const size_t b = (i * nyd + j) * nzp + k;
                Bz[b] += -dtx * (Ey[(i * nyd + j) * nzp + k] - Ey[((i - 1) * nyd + j) * nzp + k]) +
                         dty * (Ex[(i * nyp + j) * nzp + k] - Ex[(i * nyp + j - 1) * nzp + k]);
            }
}

// MA_Solver2D_norm::operator() (MA_Solver2D_norm.cpp:18-).
void ora_yee2d_ma(double *Ex, double *Ey, double *Ez,
                  const double *Bx, const double *By, const double *Bz,
                  const double *Jx, const double *Jy, const double *Jz,
                  const int np[2], double dt, const double dx[2])
{
    const double dtx = dt / dx[0], dty = dt / dx[1];
    const size_t nxp = np[0], nyp = np[1], nxd = nxp + 1, nyd = nyp + 1;
    for (size_t i = 0; i < nxd; ++i)
        for (size_t j = 0; j < nyp; ++j)
            Ex[i * nyp + j] += -dt * Jx[i * nyp + j] + dty * (Bz[i * nyd + j + 1] - Bz[i * nyd + j]);
    for (size_t i = 0; i < nxp; ++i)
        for (size_t j = 0; j < nyd; ++j)
            Ey[i * nyd + j] += -dt * Jy[i * nyd + j] - dtx * (Bz[(i + 1) * nyd + j] - Bz[i * nyd + j]);
    for (size_t i = 0; i < nxp; ++i)
        for (size_t j = 0; j < nyp; ++j)
            Ez[i * nyp + j] += -dt * Jz[i * nyp + j] + dtx * (By[(i + 1) * nyp + j] - By[i * nyp + j]) -
                               dty * (Bx[i * nyd + j + 1] - Bx[i * nyd + j]);
}

// MF_Solver2D_Yee::operator() (MF_Solver2D_Yee.cpp:18-).
void ora_yee2d_mf(double *Bx, double *By, double *Bz,
                  const double *Ex, const double *Ey, const double *Ez,
                  const int np[2], double dt, const double dx[2])
{
    const double dtx = dt / dx[0], dty = dt / dx[1];
    const size_t nxp = np[0], nyp = np[1], nxd = nxp + 1, nyd = nyp + 1;
    for (size_t i = 0; i < nxd - 1; ++i)
        for (size_t j = 1; j < nyd - 1; ++j)
            Bx[i * nyd + j] -= dty * (Ez[i * nyp + j] - Ez[i * nyp + j - 1]);
    for (size_t i = 1; i < nxd - 1; ++i)
        for (size_t j = 0; j < nyp; ++j)
            By[i * nyp + j] += dtx * (Ez[i * nyp + j] - Ez[(i - 1) * nyp + j]);
    for (size_t i = 1; i < nxd - 1; ++i)
        for (size_t j = 1; j < nyd - 1; ++j)
            Bz[i * nyd + j] += dty * (Ex[i * nyp + j] - Ex[i * nyp + j - 1]) -
                               dtx * (Ey[i * nyd + j] - Ey[(i - 1) * nyd + j]);
}

}  // extern "C"
