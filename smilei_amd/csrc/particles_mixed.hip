// particles_mixed.hip -- the mixed-precision form of the particle step.
//
// Storage per particle (split precision): the nearest grid node is an exact
// int32 (its global key, patch * cells-per-patch + cluster * 64 + node in
// the 4^3 cluster, which is also the sort key, kept in `keys`), the position
// relative to it an FP32 offset in [-1/2, 1/2) per axis, momenta u in FP32,
// weight FP32: 32 B read, 28 B written per push (+4 B weight copy of the
// out-of-place ping-pong).  Binning by the nearest node (the reference bins
// by cell, floor(x/dx)) makes the Esirkepov stencil of a particle start at
// its own node: the 64 lanes of a wavefront, which hold particles of 64
// different nodes of the cluster, then deposit at addresses that differ by
// the node alone, and their 64-bit LDS atomics meet far fewer bank
// conflicts than with the random +-1 node shift of cell binning (measured:
// 14 % of the kernel time).  The API reports cells (keys, patch counts,
// order) computed from the positions.
//
// Arithmetic: shape factors, the gather, the Boris rotation and the
// Esirkepov contributions in FP32 with fused multiply-adds; the gathered
// E/B are staged from the FP64 field arrays; each cluster's J tile is
// accumulated in LDS as 64-bit fixed point (exact integer sums) and flushed
// with FP64 atomics into the FP64 J arrays (the global J reduction stays
// FP64, as do E, B and the Yee update).  The cell change and the periodic wrap are exact integer
// operations, so no rounding can misplace a particle at a patch face.
//
// Same call chain as the FP64 kernel: Interpolator3D{2,4}Order (gather),
// PusherBoris, PartBoundCond (direction of motion), Cluster3D::CellKey,
// Projector3D{2,4}Order::currents (Esirkepov).
#include <algorithm>

#include "kernels.h"
#include "pic_device.h"

namespace smi {

namespace {

using namespace dev;

constexpr uint32_t kMask = (1u << 27) - 1;

// Quadratic / quartic B-spline weights at offsets -H..H, FP32.
template <int ORDER>
__device__ inline void splinef(float d, float *w)
{
    if constexpr (ORDER == 2) {
        w[0] = fmaf(d, fmaf(0.5f, d, -0.5f), 0.125f);
        w[1] = fmaf(-d, d, 0.75f);
        w[2] = w[0] + d;
    } else {
        const float d2 = d * d;
        // 1/384 (1-2d)^4 and friends, Horner form of Interpolator3D4Order.h:62-67
        const float a = 0.5f - d, b = 0.5f + d;
        const float a2 = a * a, b2 = b * b;
        w[0] = a2 * a2 * (1.f / 24.f);
        w[4] = b2 * b2 * (1.f / 24.f);
        w[2] = fmaf(d2, fmaf(d2, 0.25f, -0.625f), 115.f / 192.f);
        const float c = fmaf(d2, -(1.f / 6.f), 0.25f);
        w[1] = fmaf(d, fmaf(d, c, 1.f / 6.f * d2 - 11.f / 24.f), 19.f / 96.f);
        w[3] = fmaf(d, fmaf(d, c, -(1.f / 6.f) * d2 + 11.f / 24.f), 19.f / 96.f);
    }
}

// Gather Sum_{a,b,c} X[a] Y[b] Z[c] f(i+a, j+b, k+c) from an LDS tile with
// strides SX, SY (z contiguous), contracted z first, then y, then x (3 FMA
// chains).
template <int H, int SX, int SY>
__device__ inline float gatherf(const float *X, const float *Y, const float *Z, const float *t, int i, int j, int k)
{
    float r = 0.f;
#pragma unroll
    for (int a = -H; a <= H; ++a) {
        float ta = 0.f;
#pragma unroll
        for (int b = -H; b <= H; ++b) {
            const float *row = t + (i + a) * SX + (j + b) * SY + k;
            float sb = 0.f;
#pragma unroll
            for (int c = -H; c <= H; ++c) sb = fmaf(Z[c + H], row[c], sb);
            ta = fmaf(Y[b + H], sb, ta);
        }
        r = fmaf(X[a + H], ta, r);
    }
    return r;
}

// Esirkepov window in FP32 (see esirkepov_window): S0 from the old delta,
// S1 from the new delta shifted by sh in {-1,0,1}.
template <int ORDER>
__device__ inline void windowf(float dold, float dnew, int sh, float *S0w, float *DSw, int &w0)
{
    constexpr int W = ORDER + 2;
    float s[ORDER + 1], nw[ORDER + 1];
    splinef<ORDER>(dold, s);
    splinef<ORDER>(dnew, nw);
    w0 = sh < 0 ? 0 : 1;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const float a = w0 == 0 ? (m == 0 ? 0.f : s[m - 1]) : (m <= ORDER ? s[m] : 0.f);
        const float b = sh <= 0 ? (m <= ORDER ? nw[m] : 0.f) : (m >= 1 ? nw[m - 1] : 0.f);
        S0w[m] = a;
        DSw[m] = b - a;
    }
}

// FP64 B-spline weights (Horner form) of the deposit.  The deposit is done
// in FP64 from the FP32 in-cell deltas: with FP32 weights the Esirkepov
// identity sum_c DS_c T_c = S1S1S1 - S0S0S0 holds only to FP32 rounding per
// particle and step, and the Gauss-law residual random-walks (measured:
// 1.3e-6 of max|rho| after 1000 steps at order 4); in FP64 it holds to
// 1e-16, so div E - rho stays at its initial value.  FP64 VALU issue rate
// equals the FP32 one on gfx950.
template <int ORDER>
__device__ inline void splined(double d, double *w)
{
    if constexpr (ORDER == 2) {
        w[0] = fma(d, fma(0.5, d, -0.5), 0.125);
        w[1] = fma(-d, d, 0.75);
        w[2] = w[0] + d;
    } else {
        const double d2 = d * d;
        const double a = 0.5 - d, b = 0.5 + d;
        const double a2 = a * a, b2 = b * b;
        w[0] = a2 * a2 * (1.0 / 24.0);
        w[4] = b2 * b2 * (1.0 / 24.0);
        w[2] = fma(d2, fma(d2, 0.25, -0.625), 115.0 / 192.0);
        const double c = fma(d2, -(1.0 / 6.0), 0.25);
        w[1] = fma(d, fma(d, c, 1.0 / 6.0 * d2 - 11.0 / 24.0), 19.0 / 96.0);
        w[3] = fma(d, fma(d, c, -(1.0 / 6.0) * d2 + 11.0 / 24.0), 19.0 / 96.0);
    }
}

// Esirkepov window in FP64 (see esirkepov_window): S0 from the old delta,
// S1 from the new delta shifted by sh in {-1,0,1}.
template <int ORDER>
__device__ inline void windowd(double dold, double dnew, int sh, double *S0w, double *DSw, int &w0)
{
    constexpr int W = ORDER + 2;
    double s[ORDER + 1], nw[ORDER + 1];
    splined<ORDER>(dold, s);
    splined<ORDER>(dnew, nw);
    w0 = sh < 0 ? 0 : 1;
#pragma unroll
    for (int m = 0; m < W; ++m) {
        const double a = w0 == 0 ? (m == 0 ? 0. : s[m - 1]) : (m <= ORDER ? s[m] : 0.);
        const double b = sh <= 0 ? (m <= ORDER ? nw[m] : 0.) : (m >= 1 ? nw[m - 1] : 0.);
        S0w[m] = a;
        DSw[m] = b - a;
    }
}

// Fixed-point accumulation of the J tile.  A contribution c, scaled to
// units of the cluster's quantum, is rounded to an integer k in an FP64
// number offset by 1.5*2^52 (whose mantissa then holds k); the raw bits of
// that double, 0x4338000000000000 + k, are summed with an integer LDS atomic.
// The constant only reaches bits 48..63, so the low 48 bits of a tile word
// hold sum(k) exactly while |sum(k)| < 2^47.  The quantum of a cluster with
// N particles is 2^-46 * N * cr_max (cr_max = the largest contribution of
// any species of the launch): every particle adds at most once to a tile
// word and |k| <= 2^46 / N, so the sum stays in range for any N, and the
// quantum is 2^-36 of cr_max for a 1024-particle cluster.  Integer adds make
// the tile sums exact and independent of the order of the atomics.
constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52
constexpr double kUnit = 1.4210854715202004e-14;  // 2^-46

__device__ inline double fixed_decode(unsigned long long v, double quantum)
{
    const long long k = ((long long)(v << 16)) >> 16;  // sign-extend the low 48 bits
    return (double)k * quantum;
}

// Windowed 3D Esirkepov with W(b,c) = S0b (S0c + DSc/2) + DSb (S0c/2 + DSc/3)
// (the reference's four-term weight, factorised, with the current factor
// crq folded into the second row), FP64 weights and prefix sums along the
// current's axis, fixed-point tile.  The prefix sum starts at the rounding
// constant, so every step of it lands on an integer of the tile's quantum
// and is added as it stands (one FMA per face).  Only the non-zero part
// of the stencil is visited: NT window entries across the current (entries
// past the union of the old and new supports are exact zeros) and NA faces
// along it.  The face after the last window entry is left out: its prefix
// sum is sum(DS) over the whole window, zero up to FP64 rounding.  A
// particle whose nearest node is unchanged on every axis has NT = ORDER+1,
// NA = ORDER (the two supports coincide); otherwise NT = ORDER+2,
// NA = ORDER+1.  tile(c, i, j, k) gives the LDS word of component c at
// window indices relative to the particle's stencil origin.
// The tile word of the partner lane (lane ^ 1) of a pair of the
// pair-interleaved layout: DPP quad permutation [1,0,3,2] on both halves.
__device__ inline unsigned long long pair_word(unsigned long long v)
{
    typedef unsigned u2_t __attribute__((ext_vector_type(2)));
    const u2_t a = __builtin_bit_cast(u2_t, v);
    u2_t b;
    b.x = __builtin_amdgcn_mov_dpp(a.x, 0xB1, 0xF, 0xF, false);
    b.y = __builtin_amdgcn_mov_dpp(a.y, 0xB1, 0xF, 0xF, false);
    return __builtin_bit_cast(unsigned long long, b);
}

// PAIR: wavefront-level pre-reduction of the tile atomics.  pairing = 1 on
// the even lane of a pair of lanes whose particles have the same stencil
// (same node, both stayers): it adds its partner's word to its own (the
// fixed-point words are integers, so the sum of two is exact: the low 48
// bits of (C + k1) + (C + k2) are k1 + k2) and issues the one atomic;
// pairing = 2 on the odd lane, which issues none; 0: own atomic.
template <int ORDER, int NT, int NA, bool PAIR = false, class Tile>
__device__ inline void esirkepov3_fixed(const double (&S0)[3][ORDER + 2], const double (&DS)[3][ORDER + 2],
                                        const int *w0, const double *crq, Tile tile, int pairing = 0)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int bx = (c + 1) % 3, ex = (c + 2) % 3;
        double P[NT], Q[NT], A[NA];
#pragma unroll
        for (int m = 0; m < NT; ++m) {
            P[m] = fma(0.5, DS[ex][m], S0[ex][m]);
            Q[m] = fma(1.0 / 3.0, DS[ex][m], 0.5 * S0[ex][m]);
        }
        double ps = DS[c][0];
        A[0] = -crq[c] * ps;
#pragma unroll
        for (int r = 1; r < NA; ++r) {
            ps += DS[c][r];
            A[r] = -crq[c] * ps;
        }
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                const double t = fma(S0[bx][a], P[b], DS[bx][a] * Q[b]);
#pragma unroll
                for (int r = 1; r <= NA; ++r) {
                    const double acc = fma(A[r - 1], t, kRound);
                    int idx[3];
                    idx[c] = w0[c] + r;
                    idx[bx] = w0[bx] + a;
                    idx[ex] = w0[ex] + b;
                    unsigned long long v = (unsigned long long)__double_as_longlong(acc);
                    if constexpr (PAIR) {
                        const unsigned long long pv = pair_word(v);
                        if (pairing == 1) v += pv;
                        if (pairing != 2) atomicAdd(tile(c, idx[0], idx[1], idx[2]), v);
                    } else {
                        atomicAdd(tile(c, idx[0], idx[1], idx[2]), v);
                    }
                }
            }
    }
}

// LDS tiles of one cluster of the mixed kernel.  Field tile: GT^3 points
// per field from G0-GO, strides GSX, GSY, 1 (floats).  J tile: JT^3 words
// per component from G0-JO, strides JSX, JSY, 1 and JSC between components
// (64-bit words).  A particle binned to node l (0..3 per axis, tile
// coordinate l + JO) deposits on tile rows l + w0 + m with window entries
// m < ORDER + 2 and w0 in {0, 1} (esirkepov3_fixed): rows 0 .. 5 + 2H, so JT
// = 6 + 2H rows cover every deposit.  The x-stride of 76
// floats (not 8*8 = 64, a multiple of the LDS bank count) spreads the four
// x-planes of cells that the lanes of a wavefront cover over different
// banks (measured on the order-2 kernel: x-strides of 68, 76, 84 equal,
// 72 and 80 6 % slower).  JPAD = 1 pads the J rows by two words (a tile
// shape chosen by the autotuner: on the beam-plasma the padded rows meet
// fewer bank conflicts of the 64-bit atomics, 3 % faster).  The node-changer
// queue takes the LDS left in the budget that keeps four workgroups per CU
// (order 2).
template <int ORDER, int JPAD, int TX = 1, bool NACC = false>
struct MixTile {
    static constexpr int H = ORDER / 2;
    static constexpr int GO = H, GT = 4 + 2 * H + 2, JO = H + 1, JT = 4 + 2 * H + 2;
    // x extents of a tile of TX clusters along x
    static constexpr int GTX = 4 * TX + 2 * H + 2, JTX = 4 * TX + 2 * H + 2;
    static constexpr int GSY = GT, GSX = ORDER == 2 ? 76 : GT * GT, GF = GTX * GSX;
    // padded shape: rows of JT + 2 words and x-planes of 104 (order 2) or
    // 136 (order 4) words, the strides for which the 64 nodes of a cluster,
    // at one common stencil offset (the lanes of one atomic instruction),
    // spread evenly over the LDS banks (each 8-byte word address mod 32 hit
    // by two lanes, the fewest possible for 64 lanes; with unpadded x-planes
    // of JT * JT words, three or four).  Measured on the beam-plasma: 8 rows
    // of 10 words in x-planes of 104 (order 2) 49.0 ms, x-planes of 88 49.2,
    // unpadded rows of 8 in planes of 66 (also two lanes per bank pair at one
    // offset) 54.2, the earlier 9 rows of 10 in planes of 104 50.2.
    // (order 4, unpadded rows: x-planes of 104 words, not 10 x 10 = 100,
    // for the same spread: 13.45 against 13.77 ms on the thermal plasma)
    static constexpr int JSY = JT + 2 * JPAD, JSX = JPAD ? (ORDER == 2 ? 104 : 136) : (ORDER == 4 ? 104 : JT * JSY),
                         JSC = JTX * JSX;
    // node accumulation (NACC): per-node staging of the stayers of one
    // 256-particle round, KST entries of 32 bytes per node, and two sets of
    // per-node counters
    static constexpr int KST = 4;
    static constexpr int kStage = NACC ? KST * 64 * 32 + 2 * 64 * 4 : 0;
    // one cluster: four workgroups per CU (order 2; three with the node
    // staging); two clusters: two.  Order 4: three workgroups per CU with
    // unpadded J rows (53 KB; 168 registers), two with padded ones
    static constexpr int kBudget =
        ORDER == 2 ? (TX == 1 ? (NACC ? 53248 : 40960) : 81920) - 64 : (JPAD ? 65536 : 53248 - 64);
    static constexpr int kFixed = 6 * GF * 4 + 3 * JSC * 8 + TX * kMaxMerged * 64 * 4 + 16 + kStage;
    static constexpr int QMAX = ORDER == 2 ? 256 : 128;
    static constexpr int QFIT = (kBudget - kFixed) / 32 / 32 * 32;
    static constexpr int QCAP = QFIT < QMAX ? QFIT : QMAX;
    static_assert(QCAP >= 64, "LDS budget leaves no node-changer queue");
};

// Esirkepov deposit of one particle from its old and new in-cell deltas
// (dold, dnew in [-1/2, 1/2)), node shifts sh and the tile coordinates o3 of
// its stencil origin; NT/NA as esirkepov3_fixed.
template <int ORDER, int NT, int NA, int JPAD, int TX, bool PAIR = false>
__device__ inline void deposit_fixed(unsigned long long *sJ, const float *dold, const float *dnew, const int *sh,
                                     const int *o3, const double *crq, int pairing = 0)
{
    using TL = MixTile<ORDER, JPAD, TX>;
    double S0[3][ORDER + 2], DS[3][ORDER + 2];
    int w0[3];
    if constexpr (NT == ORDER + 1) {
        // node unchanged on every axis: both supports start at window entry 0
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double s0[ORDER + 1], s1[ORDER + 1];
            splined<ORDER>((double)dold[d], s0);
            splined<ORDER>((double)dnew[d], s1);
#pragma unroll
            for (int m = 0; m <= ORDER; ++m) {
                S0[d][m] = s0[m];
                DS[d][m] = s1[m] - s0[m];
            }
            S0[d][ORDER + 1] = 0.;
            DS[d][ORDER + 1] = 0.;
            w0[d] = 1;
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) windowd<ORDER>((double)dold[d], (double)dnew[d], sh[d], S0[d], DS[d], w0[d]);
    }
    unsigned long long *sJb = sJ + o3[0] * TL::JSX + o3[1] * TL::JSY + o3[2];
    esirkepov3_fixed<ORDER, NT, NA, PAIR>(S0, DS, w0, crq, [&](int c, int a, int b, int k) {
        return &sJb[c * TL::JSC + a * TL::JSX + b * TL::JSY + k];
    }, pairing);
}

// Node accumulation (NACC): the stayers of one round of the workgroup
// (blockDim.x consecutive slots of a species' cluster range) are staged per
// node (old and new deltas, weight) and deposited node by node: the thread
// of component C and node n sums the fixed-point words of its node's m
// staged particles in registers and adds each of the NA x NT x NT words of
// the node's stencil to the tile with one atomic, instead of one atomic per
// particle and word.  The accumulator starts at the rounding constant and
// each contribution is rounded onto it by one FMA, so that it ends at
// 1.5*2^52 + sum(k), k the integers the per-particle deposit adds (the same
// words, but where a product falls exactly half-way between two integers of
// the quantum: round-half-even then depends on the accumulator's parity).
// The contributions repeat the per-particle arithmetic (deposit_fixed with
// NT = ORDER + 1, NA = ORDER) operation for operation.
template <int ORDER, int JPAD, int C>
__device__ inline void node_sum(unsigned long long *sJ, const float4 (*st)[64][2], int n, int m, double cwd,
                                double inv_quantum, const double *dxdt)
{
    using TL = MixTile<ORDER, JPAD, 1>;
    constexpr int NT = ORDER + 1, NA = ORDER, BX = (C + 1) % 3, EX = (C + 2) % 3;
    double acc[NT][NT][NA];
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < NA; ++r) acc[a][b][r] = kRound;
    for (int e = 0; e < m; ++e) {
        const float4 p0 = st[e][n][0], p1 = st[e][n][1];
        const float dold[3] = {p0.x, p0.y, p0.z}, dnew[3] = {p1.x, p1.y, p1.z};
        double S0[3][NT], DS[3][NT];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double s0[ORDER + 1], s1[ORDER + 1];
            splined<ORDER>((double)dold[d], s0);
            splined<ORDER>((double)dnew[d], s1);
#pragma unroll
            for (int q = 0; q < NT; ++q) {
                S0[d][q] = s0[q];
                DS[d][q] = s1[q] - s0[q];
            }
        }
        const double cwq = cwd * (double)p0.w * inv_quantum;
        const double crq = cwq * dxdt[C];
        double P[NT], Q[NT], A[NA];
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            P[q] = fma(0.5, DS[EX][q], S0[EX][q]);
            Q[q] = fma(1.0 / 3.0, DS[EX][q], 0.5 * S0[EX][q]);
        }
        double ps = DS[C][0];
        A[0] = -crq * ps;
#pragma unroll
        for (int r = 1; r < NA; ++r) {
            ps += DS[C][r];
            A[r] = -crq * ps;
        }
#pragma unroll
        for (int a = 0; a < NT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                const double t = fma(S0[BX][a], P[b], DS[BX][a] * Q[b]);
#pragma unroll
                for (int r = 0; r < NA; ++r) acc[a][b][r] = fma(A[r], t, acc[a][b][r]);
            }
    }
    if (m == 0) return;
    // the node's stencil origin: window entry 0 at tile offset 1 on every axis
    unsigned long long *sJb = sJ + C * TL::JSC + (n >> 4) * TL::JSX + ((n >> 2) & 3) * TL::JSY + (n & 3);
#pragma unroll
    for (int a = 0; a < NT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b)
#pragma unroll
            for (int r = 0; r < NA; ++r) {
                int idx[3];
                idx[C] = 2 + r;
                idx[BX] = 1 + a;
                idx[EX] = 1 + b;
                atomicAdd(&sJb[idx[0] * TL::JSX + idx[1] * TL::JSY + idx[2]],
                          (unsigned long long)__double_as_longlong(acc[a][b][r]));
            }
}

// A tile of TX clusters adjacent along x in one patch (TX = 1: one
// cluster).  Tiles of a patch are numbered like its clusters with the x
// index divided by TX.
template <int TX>
struct Cl {
    int p, pc[3], G0[3], lc0[3];  // patch, its coordinates, first global / in-patch cell of the tile
    int64_t gbase[TX];            // first global cell key of each cluster
};

template <int TX>
__device__ inline Cl<TX> tile_of(const Geom &g, int b)
{
    Cl<TX> c;
    const int tpp = g.clpp / TX;
    c.p = b / tpp;
    const int t = b % tpp;
    const int cc[3] = {t / (g.ncl[1] * g.ncl[2]) * TX, (t / g.ncl[2]) % g.ncl[1], t % g.ncl[2]};
    c.pc[0] = c.p / (g.npatch[1] * g.npatch[2]);
    c.pc[1] = (c.p / g.npatch[2]) % g.npatch[1];
    c.pc[2] = c.p % g.npatch[2];
    for (int d = 0; d < 3; ++d) {
        c.lc0[d] = 4 * cc[d];
        c.G0[d] = c.pc[d] * g.n[d] + c.lc0[d];
    }
#pragma unroll
    for (int k = 0; k < TX; ++k)
        c.gbase[k] = (int64_t)c.p * g.cpp + (int64_t)(((cc[0] + k) * g.ncl[1] + cc[1]) * g.ncl[2] + cc[2]) * 64;
    return c;
}

// One workgroup runs the particles of every species of the launch
// (specs[0..ns)) of the tile `tile` (TX 4^3 clusters adjacent along x)
// against one field tile and one J tile.  Uniform across the workgroup;
// returns at once for an empty tile.  Thread t takes the slots t,
// t + blockDim.x, ... of the tile's ranges of perm (cluster after cluster):
// a wavefront reads 64 consecutive slots, whose particles lie in different
// cells when the sort stored the range in the lane-interleaved layout
// (cluster_slot).
template <int ORDER, int JPAD, int TX, bool PAIRK = false, bool NACC = false>
__device__ inline void mixed_cluster(const Geom &g, const MixSpecies *__restrict__ specs, int ns, double cr_max,
                                     const GatherFields &F, const JPtrs &J, int *__restrict__ err, int tile)
{
    static_assert(!NACC || (TX == 1 && !PAIRK), "node accumulation runs one-cluster tiles");
    using TL = MixTile<ORDER, JPAD, TX, NACC>;
    constexpr int H = TL::H, GO = TL::GO, GT = TL::GT, JO = TL::JO, JT = TL::JT;
    constexpr int JTX = TL::JTX;
    constexpr int GSX = TL::GSX, GSY = TL::GSY, GF = TL::GF;
    constexpr int GT3 = TL::GTX * GT * GT, JT3 = JTX * JT * JT;
    __shared__ float sF[6 * GF];
    static_assert(TL::JSC % 2 == 0, "the J tile is cleared 16 bytes at a time");
    __shared__ __align__(16) unsigned long long sJ[3 * TL::JSC];
    __shared__ int sHist[TX][kMaxMerged][64];
    // particles whose node changes on some axis: their deposit (the wider
    // stencil) is deferred to a second pass over this queue, so that the
    // wavefronts of the main loop run the narrow stencil with all lanes;
    // entry = old and new deltas, weight, packed stencil origin, shifts and
    // species
    constexpr int QCAP = TL::QCAP;
    __shared__ float sQ[QCAP][8];
    __shared__ int sQn;
    // NACC: staged stayers of the current round per node (entry e of node n:
    // old deltas and weight, new deltas) and their counts, two sets of
    // counters used in alternate rounds (one is cleared while the other is
    // read)
    constexpr int KS = NACC ? TL::KST : 1, NN = NACC ? 64 : 1;
    __shared__ float4 sStage[KS][NN][2];
    __shared__ int sCnt[2][NN];
    int par = 0;

    const Cl<TX> cl = tile_of<TX>(g, tile);
    int total = 0;
    for (int q = 0; q < ns; ++q)
#pragma unroll
        for (int k = 0; k < TX; ++k) total += specs[q].cell_off[cl.gbase[k] + 64] - specs[q].cell_off[cl.gbase[k]];
    if (total == 0) return;

    const double *src[6] = {F.Ex, F.Ey, F.Ez, F.Bx, F.By, F.Bz};
    for (int t = threadIdx.x; t < GT3; t += blockDim.x) {
        const int a = t / (GT * GT), b = (t / GT) % GT, c = t % GT;
        // 32-bit index: the engine holds fewer than 2^27 cells
        const int idx = (wrapi(cl.G0[0] - GO + a, g.N[0]) * g.N[1] + wrapi(cl.G0[1] - GO + b, g.N[1])) * g.N[2] +
                        wrapi(cl.G0[2] - GO + c, g.N[2]);
#pragma unroll
        for (int f = 0; f < 6; ++f) sF[f * GF + a * GSX + b * GSY + c] = (float)src[f][idx];
    }
    for (int t = threadIdx.x; t < 3 * TL::JSC / 2; t += blockDim.x) reinterpret_cast<ulonglong2 *>(sJ)[t] = ulonglong2{0ull, 0ull};
    for (int t = threadIdx.x; t < TX * kMaxMerged * 64; t += blockDim.x) (&sHist[0][0][0])[t] = 0;
    if constexpr (NACC)
        for (int t = threadIdx.x; t < 2 * NN; t += blockDim.x) (&sCnt[0][0])[t] = 0;
    if (threadIdx.x == 0) sQn = 0;
    __syncthreads();

    const double quantum = kUnit * (double)total * cr_max, inv_quantum = 1.0 / quantum;
    // dt/dx as an FP32 pair (hi + lo): the move applies both, so that the
    // rounding of the constant does not drift every particle the same way
    float dtdx[3], dtdx_lo[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = g.dt * g.dinv[d];
        dtdx[d] = (float)v;
        dtdx_lo[d] = (float)(v - (double)dtdx[d]);
    }
    // cr_c = q w / V * dx_c / dt with one common factor: the ratio between
    // components must be exact for the discrete continuity equation
    const double dxdt[3] = {g.dxdt[0], g.dxdt[1], g.dxdt[2]};
    for (int sid = 0; sid < ns; ++sid) {
    const MixSpecies &S = specs[sid];
    const SpeciesKM &sp = S.k;
    const GMix A = glb(S.A), B = glb(S.B);
    const gptr<const uint32_t> keysA = glb(S.keysA);
    const gptr<uint32_t> keysB = glb(S.keysB);
    const gptr<const int32_t> perm = glb(S.perm);
    int32_t *__restrict__ counts = S.counts;
    const gptr<int32_t> rank = glb(S.rank);
    const bool stay = S.stay != nullptr;
    const int lay = S.lay;
    // the species' range of each cluster of the tile, one after the other
    int s0[TX], nk[TX], N = 0;
#pragma unroll
    for (int k = 0; k < TX; ++k) {
        s0[k] = S.cell_off[cl.gbase[k]];
        nk[k] = S.cell_off[cl.gbase[k] + 64] - s0[k];
        N += nk[k];
    }
    auto particle = [&](const int rel0) {
        // cluster k of the tile and the position rel in its range
        int k = 0, rel = rel0;
        if constexpr (TX == 2) {
            k = rel0 >= nk[0] ? 1 : 0;
            rel = rel0 - (k ? nk[0] : 0);
        }
        const int s = TX == 2 && k ? s0[TX - 1] : s0[0];
        const int64_t gbase = TX == 2 && k ? cl.gbase[TX - 1] : cl.gbase[0];
        const int jslot = s + rel;
        const int i = perm[jslot];
        const int lc = (int)((int64_t)(keysA[i] & kMask) - gbase);
        float o[3] = {A.ox[i], A.oy[i], A.oz[i]};
        float u[3] = {A.ux[i], A.uy[i], A.uz[i]};
        const float w = A.w[i];
        if (lc < 0 || lc >= 64) {  // not in this cluster: corrupt state
            atomicOr(err, 1);
            keysB[jslot] = ((stay ? kPlaceAtomic : 13u) << 27) | (uint32_t)gbase;
            B.ox[jslot] = o[0]; B.oy[jslot] = o[1]; B.oz[jslot] = o[2];
            B.ux[jslot] = u[0]; B.uy[jslot] = u[1]; B.uz[jslot] = u[2];
            B.w[jslot] = w;
            atomicAdd(&counts[gbase], 1);
            return;
        }
        // node in the tile (x: the nodes of the tile's cluster k start at 4k)
        const int l[3] = {(lc >> 4) + 4 * k, (lc >> 2) & 3, lc & 3};

        // ---- shape factors and gather.  Primal stencil centred on the node;
        // the nearest dual node (global index g at (g - 1/2) dx) is l+1 for a
        // non-negative offset o, l for a negative one, at offset o -+ 1/2.
        float cp[3][5], cd[3][5];
        int dn[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const bool up = o[d] >= 0.f;
            dn[d] = l[d] + (up ? 1 : 0);
            splinef<ORDER>(o[d], cp[d]);
            splinef<ORDER>(o[d] + (up ? -0.5f : 0.5f), cd[d]);
        }
        // tile coordinates (tile origin G0-GO)
        const int p0 = l[0] + GO, p1 = l[1] + GO, p2 = l[2] + GO;
        const int d0 = dn[0] + GO, d1 = dn[1] + GO, d2 = dn[2] + GO;
        float E[3], Bf[3];
        E[0] = gatherf<H, GSX, GSY>(cd[0], cp[1], cp[2], sF + 0 * GF, d0, p1, p2);
        E[1] = gatherf<H, GSX, GSY>(cp[0], cd[1], cp[2], sF + 1 * GF, p0, d1, p2);
        E[2] = gatherf<H, GSX, GSY>(cp[0], cp[1], cd[2], sF + 2 * GF, p0, p1, d2);
        Bf[0] = gatherf<H, GSX, GSY>(cp[0], cd[1], cd[2], sF + 3 * GF, p0, d1, d2);
        Bf[1] = gatherf<H, GSX, GSY>(cd[0], cp[1], cd[2], sF + 4 * GF, d0, p1, d2);
        Bf[2] = gatherf<H, GSX, GSY>(cd[0], cd[1], cp[2], sF + 5 * GF, d0, d1, p2);

        // ---- Boris push, FP32.  gamma from 1 + |u|^2 with the squares summed
        // before adding 1 and one Newton step on rsqrt.  The electric kicks
        // apply q dt / 2m as an FP32 pair (hi + lo): the rounding of the
        // constant would otherwise bias every kick of the species the same
        // way (a systematic error, where the arithmetic's rounding is random).
        const float qmh = sp.qmh, qmh_lo = sp.qmh_lo;
        float um[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) um[d] = fmaf(qmh, E[d], fmaf(qmh_lo, E[d], u[d]));
        float g2 = 1.f + fmaf(um[0], um[0], fmaf(um[1], um[1], um[2] * um[2]));
        float ig = rsqrtf(g2);
        ig = ig * fmaf(-0.5f * g2 * ig, ig, 1.5f);
        float t[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) t[d] = qmh * ig * Bf[d];
        const float t2 = fmaf(t[0], t[0], fmaf(t[1], t[1], t[2] * t[2]));
        const float den = 1.f + t2;
        float rd = __builtin_amdgcn_rcpf(den);
        rd = fmaf(rd, fmaf(-den, rd, 1.f), rd);
        const float sfac = 2.f * rd;
        const float up[3] = {fmaf(um[1], t[2], fmaf(-um[2], t[1], um[0])), fmaf(um[2], t[0], fmaf(-um[0], t[2], um[1])),
                             fmaf(um[0], t[1], fmaf(-um[1], t[0], um[2]))};
        const float sv[3] = {sfac * t[0], sfac * t[1], sfac * t[2]};
        float un[3] = {fmaf(up[1], sv[2], fmaf(-up[2], sv[1], um[0])), fmaf(up[2], sv[0], fmaf(-up[0], sv[2], um[1])),
                       fmaf(up[0], sv[1], fmaf(-up[1], sv[0], um[2]))};
#pragma unroll
        for (int d = 0; d < 3; ++d) un[d] = fmaf(qmh, E[d], fmaf(qmh_lo, E[d], un[d]));
        g2 = 1.f + fmaf(un[0], un[0], fmaf(un[1], un[1], un[2] * un[2]));
        ig = rsqrtf(g2);
        ig = ig * fmaf(-0.5f * g2 * ig, ig, 1.5f);

        // ---- move: offsets in grid units, exact integer node change (v -+ 1
        // is exact for v in [1/2, 3/2) resp. [-3/2, -1/2)); the destination
        // patch (periodic) and in-patch node without division
        float on[3];
        int step[3], mdir[3], dpc[3], lp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float vel = un[d] * ig;
            float v = fmaf(vel, dtdx[d], fmaf(vel, dtdx_lo[d], o[d]));
            int st = 0;
            if (v < -0.5f) {
                v += 1.f;
                st = -1;
            } else if (v >= 0.5f) {
                v -= 1.f;
                st = 1;
            }
            on[d] = v;
            step[d] = st;
            lp[d] = cl.lc0[d] + l[d] + st;  // in-patch node, may leave the patch by one
            mdir[d] = lp[d] < 0 ? -1 : (lp[d] >= g.n[d] ? 1 : 0);
            lp[d] -= mdir[d] * g.n[d];
            dpc[d] = cl.pc[d] + mdir[d];
            dpc[d] = dpc[d] < 0 ? g.npatch[d] - 1 : (dpc[d] >= g.npatch[d] ? 0 : dpc[d]);
        }

        // ---- Esirkepov deposit into the FP32 LDS tile
        {
            int sh[3], o3[3], moved = 0;
            float dold[3], dnew[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                sh[d] = step[d];  // new node - old node
                moved |= sh[d];
                dold[d] = o[d];
                dnew[d] = on[d];
                o3[d] = l[d];  // old node relative to G0 == tile coordinate of stencil index 0 (JO = H+1)
            }
            // contributions in units of the cluster's fixed-point quantum
            const double cwq = sp.cwd * (double)w * inv_quantum;
            const double crq[3] = {cwq * dxdt[0], cwq * dxdt[1], cwq * dxdt[2]};
            if constexpr (NACC) {
                // stayers staged for the node round; past the staging
                // capacity of the node, deposited at once
                if (moved == 0) {
                    const int e = atomicAdd(&sCnt[par][lc], 1);
                    if (e < KS) {
                        sStage[e][lc][0] = make_float4(dold[0], dold[1], dold[2], w);
                        sStage[e][lc][1] = make_float4(dnew[0], dnew[1], dnew[2], 0.f);
                    } else {
                        deposit_fixed<ORDER, ORDER + 1, ORDER, JPAD, TX>(sJ, dold, dnew, sh, o3, crq);
                    }
                }
            } else if constexpr (PAIRK) {
                // pair-interleaved layout (the kernel is launched for it):
                // pair the lanes 2c, 2c+1 whose particles are stayers of the
                // same node (both lanes active here; the partner's value is
                // only used when its lane is live)
                int pairing = 0;
                const uint64_t live = __ballot(1);
                const int lane = threadIdx.x & 63;
                const int me = moved == 0 ? (o3[0] | (o3[1] << 4) | (o3[2] << 8)) : -1;
                const int other = __builtin_amdgcn_mov_dpp(me, 0xB1, 0xF, 0xF, false);
                if (((live >> (lane ^ 1)) & 1ull) && me >= 0 && other == me) pairing = (lane & 1) ? 2 : 1;
                if (moved == 0)
                    deposit_fixed<ORDER, ORDER + 1, ORDER, JPAD, TX, true>(sJ, dold, dnew, sh, o3, crq, pairing);
            } else {
                if (moved == 0) deposit_fixed<ORDER, ORDER + 1, ORDER, JPAD, TX>(sJ, dold, dnew, sh, o3, crq);
            }
            // node changers: one queue slot range per wavefront
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
                        const int packed = o3[0] | (o3[1] << 3) | (o3[2] << 6) | ((sh[0] + 1) << 9) |
                                           ((sh[1] + 1) << 11) | ((sh[2] + 1) << 13) | (sid << 15);
                        sQ[pos][0] = dold[0];
                        sQ[pos][1] = dold[1];
                        sQ[pos][2] = dold[2];
                        sQ[pos][3] = dnew[0];
                        sQ[pos][4] = dnew[1];
                        sQ[pos][5] = dnew[2];
                        sQ[pos][6] = w;
                        sQ[pos][7] = __int_as_float(packed);
                    } else {
                        deposit_fixed<ORDER, ORDER + 2, ORDER + 1, JPAD, TX>(sJ, dold, dnew, sh, o3, crq);
                    }
                }
            }
        }

        // ---- destination key, direction of motion, store
        const uint32_t dp = __umul24(__umul24(dpc[0], g.npatch[1]) + dpc[1], g.npatch[2]) + dpc[2];
        const uint32_t dcl = __umul24(__umul24(lp[0] >> 2, g.ncl[1]) + (lp[1] >> 2), g.ncl[2]) + (lp[2] >> 2);
        const uint32_t gkey = dp * (uint32_t)g.cpp + dcl * 64 + 16 * (lp[0] & 3) + 4 * (lp[1] & 3) + (lp[2] & 3);
        // top key bits: the direction of motion, or with `stay` the rank
        // among this cluster's stayers of the cell (placed by the sort)
        uint32_t tag = (uint32_t)((mdir[0] + 1) * 9 + (mdir[1] + 1) * 3 + (mdir[2] + 1));
        const uint32_t rc = gkey - (uint32_t)gbase;
        if (rc < 64) {
            if (stay) {
                const int r = atomicAdd(&sHist[k][sid][rc], 1);
                tag = r < (int)kPlaceMax ? (uint32_t)r : kPlaceAtomic;
            } else {
                atomicAdd(&sHist[k][sid][rc], 1);
            }
        } else {
            atomicAdd(&counts[gkey], 1);
            if (stay) tag = kPlaceAtomic;
            // cell-major position of this input particle, the tie-break of
            // the canonical in-cell order at the next sort: only read for
            // the particles that leave their cluster (the sort computes it
            // for the others)
            if (rank) {
                // (by the storage's layout: the pair-interleaved layout also
                // reaches the order-4 and persistent kernels, which do not
                // pre-reduce)
                const int nc = TX == 2 && k ? nk[TX - 1] : nk[0];
                rank[jslot] = s + layout_rank(lay, rel, nc);
            }
        }
        keysB[jslot] = (tag << 27) | gkey;
        B.ox[jslot] = on[0];
        B.oy[jslot] = on[1];
        B.oz[jslot] = on[2];
        B.ux[jslot] = un[0];
        B.uy[jslot] = un[1];
        B.uz[jslot] = un[2];
        B.w[jslot] = w;
    };
    if constexpr (NACC) {
        // rounds of blockDim.x slots, each followed by the node deposit of
        // its staged stayers (threads 0..191: component tid / 64, node
        // tid % 64; threads 192..255 clear the other counter set)
        for (int base = 0; base < N; base += blockDim.x) {
            if (base + (int)threadIdx.x < N) particle(base + (int)threadIdx.x);
            __syncthreads();
            const int c = threadIdx.x >> 6, n = threadIdx.x & 63;
            if (c < 3) {
                const int cnt = sCnt[par][n], m = cnt < KS ? cnt : KS;
                if (c == 0) node_sum<ORDER, JPAD, 0>(sJ, sStage, n, m, sp.cwd, inv_quantum, dxdt);
                else if (c == 1) node_sum<ORDER, JPAD, 1>(sJ, sStage, n, m, sp.cwd, inv_quantum, dxdt);
                else node_sum<ORDER, JPAD, 2>(sJ, sStage, n, m, sp.cwd, inv_quantum, dxdt);
            } else if (c == 3) {
                sCnt[par ^ 1][n] = 0;
            }
            __syncthreads();
            par ^= 1;
        }
    } else {
        for (int rel0 = threadIdx.x; rel0 < N; rel0 += blockDim.x) particle(rel0);
    }
    }  // species

    __syncthreads();
    for (int t = threadIdx.x; t < TX * ns * 64; t += blockDim.x) {
        const int k = t / (ns * 64), q = (t / 64) % ns, cell = t % 64;
        const int c = sHist[k][q][cell];
        if (c) {
            const MixSpecies &S = specs[q];
            const int64_t gb = TX == 2 && k ? cl.gbase[TX - 1] : cl.gbase[0];
            atomicAdd(&S.counts[gb + cell], c);
            if (S.stay) S.stay[gb + cell] = (uint8_t)(c < (int)kPlaceMax ? c : (int)kPlaceMax);
        }
    }
    // deferred deposits of the node changers
    const int nq = sQn < QCAP ? sQn : QCAP;
    for (int t = threadIdx.x; t < nq; t += blockDim.x) {
        const float dold[3] = {sQ[t][0], sQ[t][1], sQ[t][2]};
        const float dnew[3] = {sQ[t][3], sQ[t][4], sQ[t][5]};
        const int packed_ = __float_as_int(sQ[t][7]);
        const double cwq = specs[packed_ >> 15].k.cwd * (double)sQ[t][6] * inv_quantum;
        const double crq[3] = {cwq * dxdt[0], cwq * dxdt[1], cwq * dxdt[2]};
        const int packed = __float_as_int(sQ[t][7]);
        const int o3[3] = {packed & 7, (packed >> 3) & 7, (packed >> 6) & 7};
        const int sh[3] = {((packed >> 9) & 3) - 1, ((packed >> 11) & 3) - 1, ((packed >> 13) & 3) - 1};
        deposit_fixed<ORDER, ORDER + 2, ORDER + 1, JPAD, TX>(sJ, dold, dnew, sh, o3, crq);
    }
    __syncthreads();
    // flush: one tile position per thread and pass, its grid index computed
    // once for the three components (32-bit: the engine holds fewer than
    // 2^27 cells)
    const gptr<double> dstJ[3] = {glb(J.Jx), glb(J.Jy), glb(J.Jz)};
    for (int r = threadIdx.x; r < JT3; r += blockDim.x) {
        const int a = r / (JT * JT), b = (r / JT) % JT, k = r % JT;  // a < JTX
        const int gi = wrapi(cl.G0[0] - JO + a, g.N[0]);
        const int gj = wrapi(cl.G0[1] - JO + b, g.N[1]);
        const int gk = wrapi(cl.G0[2] - JO + k, g.N[2]);
        const int idx = (gi * g.N[1] + gj) * g.N[2] + gk;
        const unsigned long long *w = &sJ[a * TL::JSX + b * TL::JSY + k];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = fixed_decode(w[c * TL::JSC], quantum);
            if (v != 0.) unsafeAtomicAdd((double *)(dstJ[c] + idx), v);
        }
    }
}

// One workgroup per tile of TX clusters of the grid (xcd: XCD-aware tile
// order).  Order 2: at most 128 VGPRs, and 40 KB (one cluster) or 80 KB (two)
// of LDS, so that four wavefronts per SIMD are resident (four 256-thread or
// two 512-thread workgroups per CU).  Order 4: at most 168 VGPRs, three
// wavefronts per SIMD with the unpadded J tile (measured on the thermal
// plasma: 13.85 against 14.15 ms at two wavefronts).
template <int ORDER, int JPAD, int TX, bool PAIRK = false, bool NACC = false>
__global__ void __launch_bounds__(512, NACC ? 3 : (ORDER == 2 ? 4 : 3))
k_particles_mixed(Geom g, const MixSpecies *__restrict__ specs, int ns, double cr_max, GatherFields F, JPtrs J,
                  int *__restrict__ err, int xcd)
{
    const int b = xcd ? xcd_cluster(blockIdx.x, gridDim.x) : blockIdx.x;
    mixed_cluster<ORDER, JPAD, TX, PAIRK, NACC>(g, specs, ns, cr_max, F, J, err, b);
}

// Persistent form: a grid sized to the resident workgroups pulls clusters
// from the work list (the non-empty clusters, wcount[0] of them) through the
// queue head wcount[1]; every workgroup leaves once the list is exhausted.
template <int ORDER, int JPAD>
__global__ void __launch_bounds__(512, ORDER == 2 ? 3 : 2)
k_particles_mixed_list(Geom g, const MixSpecies *__restrict__ specs, int ns, double cr_max, GatherFields F, JPtrs J,
                       int *__restrict__ err, const int *__restrict__ work, int *__restrict__ wcount)
{
    __shared__ int sItem;
    const int nwork = wcount[0];
    for (;;) {
        __syncthreads();  // the previous cluster is done with the LDS tiles
        if (threadIdx.x == 0) sItem = atomicAdd(&wcount[1], 1);
        __syncthreads();
        const int item = sItem;
        if (item >= nwork) break;
        mixed_cluster<ORDER, JPAD, 1>(g, specs, ns, cr_max, F, J, err, work[item]);
    }
}

// Work list of the clusters holding particles of any species of the launch
// (order of the list is irrelevant).  wcount[0..1] must be zero.
__global__ void k_worklist(Geom g, const MixSpecies *__restrict__ specs, int ns, int *__restrict__ work,
                           int *__restrict__ wcount)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 0;
    if (c < g.npatches * g.clpp) {
        const Cl<1> cl = tile_of<1>(g, c);
        for (int q = 0; q < ns; ++q) total += specs[q].cell_off[cl.gbase[0] + 64] - specs[q].cell_off[cl.gbase[0]];
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

// FP64 global positions -> (nearest-node key, FP32 offsets); FP32
// momenta/weight.
__global__ void k_to_mixed(Geom g, PartPtrs X, int64_t n, MixPtrs M, uint32_t *__restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x[3] = {X.x[i], X.y[i], X.z[i]};
    int c[3];
    float off[3];
    for (int d = 0; d < 3; ++d) split_node(x[d] * g.dinv[d], g.N[d], c[d], off[d]);
    keys[i] = (13u << 27) | cluster_key3(g, c);
    M.ox[i] = off[0];
    M.oy[i] = off[1];
    M.oz[i] = off[2];
    M.ux[i] = (float)X.px[i];
    M.uy[i] = (float)X.py[i];
    M.uz[i] = (float)X.pz[i];
    M.w[i] = (float)X.w[i];
}

// Global node of the particle with key gk (cluster-major key of the node).
__device__ inline void node_of(const Geom &g, uint32_t gk, int *c)
{
    const int p = (int)(gk / (uint32_t)g.cpp), key = (int)(gk % (uint32_t)g.cpp);
    const int pc[3] = {p / (g.npatch[1] * g.npatch[2]), (p / g.npatch[2]) % g.npatch[1], p % g.npatch[2]};
    const int cl = key >> 6, lc = key & 63;
    const int cc[3] = {cl / (g.ncl[1] * g.ncl[2]), (cl / g.ncl[2]) % g.ncl[1], cl % g.ncl[2]};
    const int l[3] = {lc >> 4, (lc >> 2) & 3, lc & 3};
    for (int d = 0; d < 3; ++d) c[d] = pc[d] * g.n[d] + 4 * cc[d] + l[d];
}

// Buffer slot i -> FP64 global position (x = (node + offset) * dx, wrapped
// into the box) and the key of the cell it lies in (direction 13).
__global__ void k_from_mixed(Geom g, MixPtrs M, const uint32_t *__restrict__ keys, int64_t n, PartPtrs X,
                             uint32_t *__restrict__ ckeys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3];
    node_of(g, keys[i] & kMask, c);
    const float off[3] = {M.ox[i], M.oy[i], M.oz[i]};
    double xo[3];
    int cell[3];
    for (int d = 0; d < 3; ++d) {
        double v = ((double)c[d] + (double)off[d]) * g.dx[d];
        if (v < 0.) v += g.xmax[d];
        xo[d] = v;
        cell[d] = wrapi(c[d] - (off[d] < 0.f ? 1 : 0), g.N[d]);
    }
    X.x[i] = xo[0];
    X.y[i] = xo[1];
    X.z[i] = xo[2];
    X.px[i] = M.ux[i];
    X.py[i] = M.uy[i];
    X.pz[i] = M.uz[i];
    X.w[i] = M.w[i];
    ckeys[i] = (13u << 27) | cluster_key3(g, cell);
}

// Particles per patch, by the cell each position lies in.
__global__ void k_patch_counts_mixed(Geom g, MixPtrs M, const uint32_t *__restrict__ keys, int64_t n,
                                     int32_t *__restrict__ counts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3];
    node_of(g, keys[i] & kMask, c);
    const float off[3] = {M.ox[i], M.oy[i], M.oz[i]};
    int pc[3];
    for (int d = 0; d < 3; ++d) pc[d] = wrapi(c[d] - (off[d] < 0.f ? 1 : 0), g.N[d]) / g.n[d];
    atomicAdd(&counts[(pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2]], 1);
}

// Kinetic energy sum w (gamma - 1) with gamma - 1 = |u|^2 / (1 + gamma)
// (no cancellation for cold particles), accumulated in FP64.
__global__ void k_ukin_mixed(MixPtrs M, const int32_t *__restrict__ perm, int64_t n, double *partials)
{
    __shared__ double part[16];
    double s = 0.;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const int i = perm[j];
        const double ux = M.ux[i], uy = M.uy[i], uz = M.uz[i];
        const double u2 = ux * ux + uy * uy + uz * uz;
        s += (double)M.w[i] * (u2 / (1.0 + sqrt(1.0 + u2)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        partials[blockIdx.x] = t;
    }
}

// Charge density of the mixed particles (FP64 accumulation): primal
// stencil centred on the particle's node.
template <int ORDER>
__global__ void k_rho_mixed(Geom g, SpeciesKM sp, MixPtrs M, const uint32_t *__restrict__ keys,
                            const int32_t *__restrict__ perm, int64_t n, double *__restrict__ rho)
{
    constexpr int H = ORDER / 2, W = 2 * H + 1;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int i = perm[j];
    int c[3];
    node_of(g, keys[i] & kMask, c);
    const float off[3] = {M.ox[i], M.oy[i], M.oz[i]};
    double S[3][5];
    int base[3];
    for (int d = 0; d < 3; ++d) {
        float w[5];
        splinef<ORDER>(off[d], w);
        for (int a = 0; a < W; ++a) S[d][a] = w[a];
        base[d] = c[d] - H;
    }
    const double cwv = sp.cwd * (double)M.w[i];
    for (int a = 0; a < W; ++a)
        for (int b = 0; b < W; ++b)
            for (int k = 0; k < W; ++k) {
                const int64_t idx = ((int64_t)wrapi(base[0] + a, g.N[0]) * g.N[1] + wrapi(base[1] + b, g.N[1])) *
                                        g.N[2] + wrapi(base[2] + k, g.N[2]);
                unsafeAtomicAdd(&rho[idx], cwv * S[0][a] * S[1][b] * S[2][k]);
            }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

bool mixed_tile_x_ok(const Geom &g, int tile_x, bool persistent)
{
    if (tile_x == 1) return true;
    // two clusters along x: order 2, an even number of clusters per patch
    // along x, grid launch
    return tile_x == 2 && g.order == 2 && g.ncl[0] % 2 == 0 && !persistent;
}

bool mixed_node_acc_ok(const Geom &g, int threads, int tile_x, bool paired, bool persistent)
{
    // order 2, one-cluster tiles, the lane or cell-major layouts, 256-thread
    // grid launch (three wavefronts sum the three components of 64 nodes)
    return g.order == 2 && tile_x == 1 && !paired && threads == 256 && !persistent;
}

void launch_particles_mixed(const Geom &g, const MixSpecies *specs, int ns, double cr_max, const GatherFields &F,
                            const JPtrs &J, int *err, int threads, bool xcd, int tile, int tile_x, bool paired,
                            bool node_acc, int persistent_blocks, int *work, int *wcount, hipStream_t st)
{
    if (tile != 0 && tile != 1) throw Error(-1, "mixed tile shape must be 0 or 1");
    if (g.nd != 3) throw Error(-1, "the mixed-precision form is implemented for 3Dcartesian");
    if (ns < 1 || ns > kMaxMerged) throw Error(-1, "bad species count of a merged particle launch");
    if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
        throw Error(-1, "particle-kernel workgroup size must be 64, 128, 256 or 512");
    if (!mixed_tile_x_ok(g, tile_x, persistent_blocks > 0))
        throw Error(-1, "patch tile of this extent not available for this geometry or launch form");
    if (node_acc && !mixed_node_acc_ok(g, threads, tile_x, paired, persistent_blocks > 0))
        throw Error(-1, "node-accumulated deposit not available for this geometry or launch form");
    const int nclusters = g.npatches * g.clpp;
    const dim3 block(threads);
    if (persistent_blocks <= 0) {
        const dim3 grid(nclusters / tile_x);
#define SMI_PM(O, P, T) hipLaunchKernelGGL((k_particles_mixed<O, P, T>), grid, block, 0, st, g, specs, ns, \
                                           cr_max, F, J, err, xcd ? 1 : 0)
        if (g.order == 2) {
            if (tile_x == 2) {
                if (tile) SMI_PM(2, 1, 2);
                else SMI_PM(2, 0, 2);
            } else if (node_acc) {
                // stayers deposited node by node (one atomic per node and word)
                if (tile) hipLaunchKernelGGL((k_particles_mixed<2, 1, 1, false, true>), grid, block, 0, st, g, specs,
                                             ns, cr_max, F, J, err, xcd ? 1 : 0);
                else hipLaunchKernelGGL((k_particles_mixed<2, 0, 1, false, true>), grid, block, 0, st, g, specs, ns,
                                        cr_max, F, J, err, xcd ? 1 : 0);
            } else if (paired) {
                // pair pre-reduction of the tile atomics (one-cluster tiles)
                if (tile) hipLaunchKernelGGL((k_particles_mixed<2, 1, 1, true>), grid, block, 0, st, g, specs, ns,
                                             cr_max, F, J, err, xcd ? 1 : 0);
                else hipLaunchKernelGGL((k_particles_mixed<2, 0, 1, true>), grid, block, 0, st, g, specs, ns,
                                        cr_max, F, J, err, xcd ? 1 : 0);
            } else {
                if (tile) SMI_PM(2, 1, 1);
                else SMI_PM(2, 0, 1);
            }
        } else if (paired) {
            if (tile) hipLaunchKernelGGL((k_particles_mixed<4, 1, 1, true>), grid, block, 0, st, g, specs, ns,
                                         cr_max, F, J, err, xcd ? 1 : 0);
            else hipLaunchKernelGGL((k_particles_mixed<4, 0, 1, true>), grid, block, 0, st, g, specs, ns, cr_max,
                                    F, J, err, xcd ? 1 : 0);
        } else {
            if (tile) SMI_PM(4, 1, 1);
            else SMI_PM(4, 0, 1);
        }
#undef SMI_PM
        SMI_LAUNCH_CHECK();
        return;
    }
    // persistent: build the work list, then as many workgroups as fit at once
    launch_fill_i32(wcount, 2, 0, st);
    hipLaunchKernelGGL(k_worklist, dim3((nclusters + 255) / 256), dim3(256), 0, st, g, specs, ns, work, wcount);
    SMI_LAUNCH_CHECK();
    const int nblocks = std::min(nclusters, persistent_blocks);
#define SMI_PL(O, P) hipLaunchKernelGGL((k_particles_mixed_list<O, P>), dim3(nblocks), block, 0, st, g, specs, \
                                        ns, cr_max, F, J, err, (const int *)work, wcount)
    if (g.order == 2) {
        if (tile) SMI_PL(2, 1);
        else SMI_PL(2, 0);
    } else {
        if (tile) SMI_PL(4, 1);
        else SMI_PL(4, 0);
    }
#undef SMI_PL
    SMI_LAUNCH_CHECK();
}

int mixed_persistent_blocks(const Geom &g, int threads, int tile)
{
    int dev = 0, cus = 0, per_cu = 0;
    SMI_HIP(hipGetDevice(&dev));
    SMI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void *fn = nullptr;
    if (g.order == 2) fn = tile ? (const void *)k_particles_mixed_list<2, 1> : (const void *)k_particles_mixed_list<2, 0>;
    else fn = tile ? (const void *)k_particles_mixed_list<4, 1> : (const void *)k_particles_mixed_list<4, 0>;
    SMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
    return std::max(1, cus * std::max(per_cu, 1));
}

void launch_to_mixed(const Geom &g, const PartPtrs &X, int64_t n, const MixPtrs &M, uint32_t *keys, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_mixed, dim3(nblk(n, 256)), dim3(256), 0, st, g, X, n, M, keys);
    SMI_LAUNCH_CHECK();
}

void launch_from_mixed(const Geom &g, const MixPtrs &M, const uint32_t *keys, int64_t n, const PartPtrs &X,
                       uint32_t *cell_keys, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_from_mixed, dim3(nblk(n, 256)), dim3(256), 0, st, g, M, keys, n, X, cell_keys);
    SMI_LAUNCH_CHECK();
}

void launch_patch_counts_mixed(const Geom &g, const MixPtrs &M, const uint32_t *keys, int64_t n, int32_t *counts,
                               hipStream_t st)
{
    launch_fill_i32(counts, g.npatches, 0, st);
    if (n == 0) return;
    hipLaunchKernelGGL(k_patch_counts_mixed, dim3(nblk(n, 256)), dim3(256), 0, st, g, M, keys, n, counts);
    SMI_LAUNCH_CHECK();
}

void launch_ukin_mixed(const MixPtrs &M, const int32_t *perm, int64_t n, double *partials, int nblocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_ukin_mixed, dim3(nblocks), dim3(256), 0, st, M, perm, n, partials);
    SMI_LAUNCH_CHECK();
}

void launch_rho_mixed(const Geom &g, const SpeciesKM &sp, const MixPtrs &M, const uint32_t *keys, const int32_t *perm,
                      int64_t n, double *rho, hipStream_t st)
{
    if (n == 0) return;
    if (g.order == 2)
        hipLaunchKernelGGL(k_rho_mixed<2>, dim3(nblk(n, 256)), dim3(256), 0, st, g, sp, M, keys, perm, n, rho);
    else
        hipLaunchKernelGGL(k_rho_mixed<4>, dim3(nblk(n, 256)), dim3(256), 0, st, g, sp, M, keys, perm, n, rho);
    SMI_LAUNCH_CHECK();
}

}  // namespace smi
