// sort.hip -- per-step cell sort and inter-patch exchange of particles.
//
// Replaces the reference's leaving-particle extraction, neighbour exchange
// and per-patch sort (Particles::copyLeavingParticlesToBuffers
// Particles.cpp:1308, Patch::exchParticles/cornersParticles Patch.cpp:
// 670-801, nvidiaParticles::sortParticleByKey nvidiaParticles.cu:1217-1247)
// with one device-wide counting sort by destination cell, since the push
// kernel already computed every particle's destination patch and cell:
//   1. histogram of destination cells          (int32 atomics)
//   2. exclusive scan -> cell_off               (3-phase block scan)
//   3. scatter of slot indices into buckets     (atomic count-down, which
//      also returns the histogram to zero for the next step)
//   4. per-bucket rank sort by (direction of motion, source slot), the
//      canonical order of particles inside a cell.
// The particle data itself is not moved here: the next step's particle
// kernel reads its input through `perm`.
#include "kernels.h"
#include "pic_device.h"

namespace smi {

namespace {

constexpr uint32_t kKeyMask = (1u << 27) - 1;

__global__ void k_hist(const uint32_t *__restrict__ keys, int64_t n, int32_t *__restrict__ counts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&counts[keys[i] & kKeyMask], 1);
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one int per thread; returns the block total.
__device__ inline int block_exclusive_scan(int v, int &total)
{
    __shared__ int warp_tot[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) warp_tot[wid] = incl;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wid) base += warp_tot[w];
        tot += warp_tot[w];
    }
    __syncthreads();
    total = tot;
    return base + incl - v;
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_reduce(const int32_t *__restrict__ in, int64_t n, int32_t *__restrict__ bsum)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
        if (i < n) s += in[i];
    }
    int tot;
    block_exclusive_scan(s, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Single-block exclusive scan of the block sums (in place).
__global__ void __launch_bounds__(kScanThreads) k_scan_bsum(int32_t *__restrict__ bsum, int64_t nb)
{
    int carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kScanTile) {
        int v[kScanItems];
        int s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = b0 + (int64_t)threadIdx.x * kScanItems + k;
            v[k] = i < nb ? bsum[i] : 0;
            s += v[k];
        }
        int tot;
        int ex = block_exclusive_scan(s, tot) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = b0 + (int64_t)threadIdx.x * kScanItems + k;
            if (i < nb) bsum[i] = ex;
            ex += v[k];
        }
        carry += tot;
    }
}

// out[i] = bsum[block] + exclusive prefix inside the block; out[n] = total.
// Each thread owns 16 consecutive entries, moved as four 16-byte vectors
// (the entries of a full tile are 64-byte aligned); the ragged last tile
// takes the scalar path.  stay (placed stayers, or null): in[i] -= stay[i]
// after the totals are read, leaving the count of the bucket's atomic
// places, and stay[i] = 0.  clus_old (or null): the previous out[] of every
// cpcell-th entry (each cluster's first slot before this sort) is saved
// there before it is overwritten.
__global__ void __launch_bounds__(kScanThreads)
k_scan_apply(int32_t *__restrict__ in, int64_t n, const int32_t *__restrict__ bsum, int32_t *__restrict__ out,
             uint8_t *__restrict__ stay, int32_t *__restrict__ clus_old, int cpcell)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    const bool full = (int64_t)(blockIdx.x + 1) * kScanTile <= n;
    if (clus_old) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n && (base + k) % cpcell == 0) clus_old[(base + k) / cpcell] = out[base + k];
    }
    int v[kScanItems];
    int s = 0;
    if (full) {
        const int4 *src = reinterpret_cast<const int4 *>(in + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const int4 t = src[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += v[k];
    int tot;
    int ex = block_exclusive_scan(s, tot) + bsum[blockIdx.x];
    int o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        o[k] = ex;
        ex += v[k];
    }
    if (full) {
        int4 *dst = reinterpret_cast<int4 *>(out + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) dst[q] = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
    // the grand total after the last entry
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k == n - 1) out[n] = o[k] + v[k];
    if (!stay) return;
    if (full) {
        uint4 *sp = reinterpret_cast<uint4 *>(stay + base);
        const uint4 t = *sp;
        const uint32_t wd[4] = {t.x, t.y, t.z, t.w};
        if (wd[0] | wd[1] | wd[2] | wd[3]) {
            int4 *dst = reinterpret_cast<int4 *>(in + base);
#pragma unroll
            for (int q = 0; q < kScanItems / 4; ++q)
                dst[q] = make_int4(v[4 * q] - (int)(wd[q] & 255u), v[4 * q + 1] - (int)((wd[q] >> 8) & 255u),
                                   v[4 * q + 2] - (int)((wd[q] >> 16) & 255u), v[4 * q + 3] - (int)(wd[q] >> 24));
            *sp = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n && stay[base + k]) {
                in[base + k] = v[k] - stay[base + k];
                stay[base + k] = 0;
            }
    }
}

// Scatter of slot indices into their cell buckets.  Consecutive slots mostly
// keep their cell, so lanes of a wavefront with equal keys form runs: the
// run head takes `run` places with one atomic and the lanes fill them in
// slot order.  The count-down returns the histogram to zero.  With the
// count c before a run's atomic, the run takes places [end - c, end - c +
// run) of its bucket [start, end): successive runs fill the top counts[g]
// places, the whole bucket without placed stayers.  placed: a key whose top
// bits hold a rank r < kPlaceAtomic goes to start + r with no atomic (its
// cluster's stayers, SortBufs::stay).  cpc > 0 (the cells per cluster): the
// cell-major place is stored at its lane-interleaved slot of the cluster
// (cluster_slot).  The slots are dealt to the workgroups in the XCD-aware
// order (xcd_cluster): consecutive slots, and so the writes into one
// cluster's range of perm, meet in one L2, where the partial lines merge
// before they are written back.  Each wavefront takes kScatItems groups of
// 64 consecutive slots and issues the loads, atomics and stores of the
// groups together: the kernel is bound by the latency of its dependent
// accesses (key, count-down, bucket offsets, store), not by bytes.  Six
// groups per wavefront measured best on the 1.08e9-particle beam-plasma
// (sort 5.13 ms per step; four groups 5.40, five 5.26, seven 5.16, eight
// 5.62).
constexpr int kScatThreads = 256, kScatItems = 6, kScatTile = kScatThreads * kScatItems;

__global__ void __launch_bounds__(kScatThreads)
k_scatter(const uint32_t *__restrict__ keys, int64_t n, const int32_t *__restrict__ cell_off,
          int32_t *__restrict__ counts, int32_t *__restrict__ perm, int cpc, int lay, int placed)
{
    const int lane = threadIdx.x & 63;
    const int64_t i0 = (int64_t)dev::xcd_cluster(blockIdx.x, gridDim.x) * kScatTile +
                       (int64_t)(threadIdx.x >> 6) * (64 * kScatItems) + lane;
    uint32_t kk[kScatItems];
#pragma unroll
    for (int j = 0; j < kScatItems; ++j) kk[j] = i0 + 64 * j < n ? keys[i0 + 64 * j] : 0u;
    int c[kScatItems], my_head[kScatItems];
#pragma unroll
    for (int j = 0; j < kScatItems; ++j) {
        const bool direct = placed && (kk[j] >> 27) < kPlaceAtomic;
        // lanes of the count-down; the others carry the sentinel
        const uint32_t g = i0 + 64 * j < n && !direct ? kk[j] & kKeyMask : 0xFFFFFFFFu;
        const uint32_t gp = __shfl_up(g, 1, 64);
        const bool head = lane == 0 || gp != g;
        const uint64_t heads = __ballot(head);
        const uint64_t above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const int next = above ? __ffsll((unsigned long long)above) - 1 : 64;
        my_head[j] = 63 - __clzll((unsigned long long)(heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1))));
        c[j] = 0;
        if (head && g != 0xFFFFFFFFu) c[j] = atomicSub(&counts[g], next - lane);
    }
#pragma unroll
    for (int j = 0; j < kScatItems; ++j) {
        const int ch = __shfl(c[j], my_head[j], 64);
        const int64_t i = i0 + 64 * j;
        if (i >= n) continue;
        const uint32_t tag = kk[j] >> 27;
        const bool direct = placed && tag < kPlaceAtomic;
        const uint32_t gc = kk[j] & kKeyMask;
        const int pos = direct ? cell_off[gc] + (int)tag : cell_off[gc + 1] - ch + (lane - my_head[j]);
        if (cpc) {
            const uint32_t cb = gc & ~(uint32_t)(cpc - 1);
            const int s = cell_off[cb];
            perm[s + layout_slot(lay, pos - s, cell_off[cb + cpc] - s)] = (int32_t)i;
        } else {
            perm[pos] = (int32_t)i;
        }
    }
}

// Canonical order inside a cell: by direction of motion, then source slot.
// One thread per cell.  Buckets of up to kFixRegs particles are ranked in
// registers (all loads issued together, rank = number of smaller
// (direction, slot) pairs; the pairs are distinct because slots are);
// buckets of up to kFixMid are left to k_bucket_fix_mid, larger ones go to
// the work list `big` (count in nbig[0], appended by k_bucket_fix_mid) for
// k_bucket_fix_big, or, without a list, to an insertion sort in memory.
// The cell-major input is src; cpc == 0: sorted in place (dst == src);
// cpc > 0: every particle is written to its lane-interleaved slot of the
// cluster in dst.  The tie-break value of source slot v is tie[v] when tie
// is given, else v.
constexpr int kFixRegs = 16;
constexpr int kFixMid = 64;

// Where the tie-breaks come from: tie[v] (the cell-major position of slot
// v in the previous sorted order, recorded by the particle kernel), or v
// itself when tie is null.  With clus_old (each cluster's first slot in
// the previous order, stored lane-interleaved), the tie of a slot inside
// the previous range of the cell's own cluster -- a particle that stayed
// in its cluster, nearly all of them -- is computed (cluster_rank) instead
// of gathered.  Such a particle did not leave its patch, so its direction of
// motion is 13 (no patch face crossed) and its key need not be read either,
// unless some axis is one cluster wide (a periodic wrap lands in the same
// cluster): stay13 = 0 then.
struct TieSrc {
    const int32_t *tie;
    const int32_t *clus_old;
    int cpcell, ncl;
    int lay;  // layout of the previous order's clusters
    int64_t n;
    int stay13;
    // previous range [cs, cs + cn) of the cluster of cell g (cn = 0: none)
    __device__ void range(int64_t g, int &cs, int &cn) const
    {
        cs = 0;
        cn = 0;
        if (!tie || !clus_old) return;
        const int b = (int)(g / cpcell);
        cs = clus_old[b];
        cn = (b + 1 < ncl ? clus_old[b + 1] : (int)n) - cs;
    }
};

__device__ inline uint64_t fix_key(const uint32_t *keys, const TieSrc &T, int32_t v, int cs, int cn)
{
    uint32_t t = (uint32_t)v;
    if (T.tie) {
        const int r = v - cs;
        if (r >= 0 && r < cn) {
            t = (uint32_t)(cs + layout_rank(T.lay, r, cn));
            if (T.stay13) return (13ull << 32) | t;
        } else {
            t = (uint32_t)T.tie[v];
        }
    }
    return ((uint64_t)(keys[v] >> 27) << 32) | t;
}

__global__ void k_bucket_fix(const uint32_t *__restrict__ keys, int64_t ncells, const int32_t *__restrict__ cell_off,
                             int32_t *src, int32_t *dst, int cpc, int lay, TieSrc T, int32_t *__restrict__ big,
                             int32_t *__restrict__ nbig)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ncells) return;
    const int a = cell_off[g], b = cell_off[g + 1];
    const int n = b - a;
    int cs = 0, cn = 0;
    if (cpc) {
        const int64_t cb = g & ~(int64_t)(cpc - 1);
        cs = cell_off[cb];
        cn = cell_off[cb + cpc] - cs;
    }
    // final place of the particle of cell-major position pos
    auto put = [&](int pos, int32_t v) { dst[cpc ? cs + layout_slot(lay, pos - cs, cn) : pos] = v; };
    if (n < 2) {
        if (cpc && n == 1) put(a, src[a]);
        return;
    }
    int ocs, ocn;
    T.range(g, ocs, ocn);
    if (n <= kFixRegs) {
        int32_t v[kFixRegs];
        uint64_t k[kFixRegs];
#pragma unroll
        for (int i = 0; i < kFixRegs; ++i) v[i] = i < n ? src[a + i] : 0;
#pragma unroll
        for (int i = 0; i < kFixRegs; ++i) k[i] = i < n ? fix_key(keys, T, v[i], ocs, ocn) : ~0ull;
        bool sorted = true;
#pragma unroll
        for (int i = 1; i < kFixRegs; ++i) sorted &= k[i - 1] < k[i] || i >= n;
        if (sorted) {
            if (cpc)
#pragma unroll
                for (int i = 0; i < kFixRegs; ++i)
                    if (i < n) put(a + i, v[i]);
            return;
        }
#pragma unroll
        for (int i = 0; i < kFixRegs; ++i) {
            int r = 0;
#pragma unroll
            for (int j = 0; j < kFixRegs; ++j) r += k[j] < k[i] ? 1 : 0;
            if (i < n) put(a + r, v[i]);
        }
        return;
    }
    if (big) return;  // k_bucket_fix_mid, k_bucket_fix_big
    int32_t *perm = src;
    for (int i = a + 1; i < b; ++i) {
        const int32_t v = perm[i];
        const uint64_t kv = fix_key(keys, T, v, ocs, ocn);
        int j = i - 1;
        while (j >= a) {
            const int32_t u = perm[j];
            const uint64_t ku = fix_key(keys, T, u, ocs, ocn);
            if (ku <= kv) break;
            perm[j + 1] = u;
            --j;
        }
        perm[j + 1] = v;
    }
    if (cpc)
        for (int i = a; i < b; ++i) put(i, perm[i]);
}

// Wave-aggregated append of the lanes with `want` to a counter: returns
// each such lane's index (one atomic per wavefront).
__device__ inline int wave_append(bool want, int32_t *counter)
{
    const uint64_t m = __ballot(want);
    if (!m) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    return base + __popcll(m & ((1ull << lane) - 1));
}

// Each wavefront scans 64 consecutive cells and runs those with kFixRegs <
// n <= kFixMid particles one after the other, with the whole wavefront:
// lane l holds place l, its rank = the number of smaller keys over the
// wavefront (shuffles).  The stores follow the loads of every lane, so the
// in-place case is safe.  Cells with more particles go to the work list
// `big` for k_bucket_fix_big.
__global__ void __launch_bounds__(256)
k_bucket_fix_mid(const uint32_t *__restrict__ keys, int64_t ncells, const int32_t *__restrict__ cell_off,
                 int32_t *src, int32_t *dst, int cpc, int lay, TieSrc T, int32_t *__restrict__ big,
                 int32_t *__restrict__ nbig)
{
    const int lane = threadIdx.x & 63;
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    const int64_t gl = g0 + lane;
    const int nl = gl < ncells ? cell_off[gl + 1] - cell_off[gl] : 0;
    const int ib = wave_append(nl > kFixMid, &nbig[0]);
    if (nl > kFixMid) big[ib] = (int32_t)gl;
    uint64_t todo = __ballot(nl > kFixRegs && nl <= kFixMid);
    while (todo) {
        const int bit = __ffsll((unsigned long long)todo) - 1;
        todo &= todo - 1;
        const int64_t g = g0 + bit;
        const int a = cell_off[g], n = cell_off[g + 1] - a;
        int cs = 0, cn = 0;
        if (cpc) {
            const int64_t cb = g & ~(int64_t)(cpc - 1);
            cs = cell_off[cb];
            cn = cell_off[cb + cpc] - cs;
        }
        int ocs, ocn;
        T.range(g, ocs, ocn);
        const int32_t v = lane < n ? src[a + lane] : 0;
        const uint64_t k = lane < n ? fix_key(keys, T, v, ocs, ocn) : ~0ull;
        int r = 0;
        for (int j = 0; j < n; ++j) r += __shfl(k, j, 64) < k ? 1 : 0;
        if (lane < n) dst[cpc ? cs + layout_slot(lay, a + r - cs, cn) : a + r] = v;
    }
}

// The cells of the work list (more than kFixMid particles; their number in
// nbig[0], nbig[1] = 0 on entry), one workgroup per cell pulled from the
// list until it is exhausted.  Up to kBigLds particles: bitonic sort of the
// (key, slot) pairs in LDS.  More: each particle's rank = the number of
// smaller keys, counted over the cell in LDS-staged chunks, kBigLds target
// particles at a time (in place, cpc == 0, the cell is first copied to scr
// so that no write can overtake a read).  Every workgroup exits once the
// list is exhausted.
constexpr int kBigThreads = 256;
constexpr int kBigLds = 4096;
constexpr int kBigPer = kBigLds / kBigThreads;
constexpr int kBigBlocks = 1024;

__global__ void __launch_bounds__(kBigThreads)
k_bucket_fix_big(const uint32_t *__restrict__ keys, const int32_t *__restrict__ cell_off, int32_t *src, int32_t *dst,
                 int32_t *scr, int cpc, int lay, TieSrc T, const int32_t *__restrict__ big,
                 int32_t *__restrict__ nbig)
{
    __shared__ uint64_t sK[kBigLds];
    __shared__ int32_t sV[kBigLds];
    __shared__ int sItem;
    const int nb = nbig[0];
    const int tid = threadIdx.x;
    for (;;) {
        __syncthreads();  // the previous cell is done with the LDS arrays
        if (tid == 0) sItem = atomicAdd(&nbig[1], 1);
        __syncthreads();
        const int item = sItem;
        if (item >= nb) break;
        const int64_t g = big[item];
        const int a = cell_off[g], n = cell_off[g + 1] - a;
        int cs = 0, cn = 0;
        if (cpc) {
            const int64_t cb = g & ~(int64_t)(cpc - 1);
            cs = cell_off[cb];
            cn = cell_off[cb + cpc] - cs;
        }
        auto put = [&](int pos, int32_t v) { dst[cpc ? cs + layout_slot(lay, pos - cs, cn) : pos] = v; };
        int ocs, ocn;
        T.range(g, ocs, ocn);
        if (n <= kBigLds) {
            int P = 1;
            while (P < n) P <<= 1;
            for (int i = tid; i < P; i += kBigThreads) {
                const int32_t v = i < n ? src[a + i] : 0;
                sK[i] = i < n ? fix_key(keys, T, v, ocs, ocn) : ~0ull;
                sV[i] = v;
            }
            __syncthreads();
            for (int k = 2; k <= P; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = tid; i < P; i += kBigThreads) {
                        const int l = i ^ j;
                        if (l > i && ((sK[i] > sK[l]) == ((i & k) == 0))) {
                            const uint64_t tk = sK[i];
                            sK[i] = sK[l];
                            sK[l] = tk;
                            const int32_t tv = sV[i];
                            sV[i] = sV[l];
                            sV[l] = tv;
                        }
                    }
                    __syncthreads();
                }
            for (int i = tid; i < n; i += kBigThreads) put(a + i, sV[i]);
            continue;
        }
        const int32_t *rd = src;
        if (!cpc) {
            for (int i = tid; i < n; i += kBigThreads) scr[a + i] = src[a + i];
            __threadfence_block();
            __syncthreads();
            rd = scr;
        }
        for (int t0 = 0; t0 < n; t0 += kBigLds) {
            uint64_t mk[kBigPer];
            int32_t mv[kBigPer];
            int rk[kBigPer];
#pragma unroll
            for (int m = 0; m < kBigPer; ++m) {
                const int i = t0 + tid + m * kBigThreads;
                mv[m] = i < n ? rd[a + i] : 0;
                mk[m] = i < n ? fix_key(keys, T, mv[m], ocs, ocn) : ~0ull;
                rk[m] = 0;
            }
            for (int c0 = 0; c0 < n; c0 += kBigLds) {
                const int cnum = n - c0 < kBigLds ? n - c0 : kBigLds;
                __syncthreads();
                for (int i = tid; i < cnum; i += kBigThreads) sK[i] = fix_key(keys, T, rd[a + c0 + i], ocs, ocn);
                __syncthreads();
                for (int j = 0; j < cnum; ++j) {
                    const uint64_t kj = sK[j];
#pragma unroll
                    for (int m = 0; m < kBigPer; ++m) rk[m] += kj < mk[m] ? 1 : 0;
                }
            }
#pragma unroll
            for (int m = 0; m < kBigPer; ++m)
                if (t0 + tid + m * kBigThreads < n) put(a + rk[m], mv[m]);
        }
    }
}

// Slot j of the current order, stored with cluster layout `from`, moved to
// its slot under layout `to` (layout_slot); the cluster of the particle
// comes from its key, cpc = cells per cluster.
__global__ void k_relayout(const uint32_t *__restrict__ keys, int64_t n, const int32_t *__restrict__ cell_off,
                           const int32_t *__restrict__ src, int32_t *__restrict__ dst, int cpc, int from, int to)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int32_t v = src[j];
    const uint32_t cb = (keys[v] & kKeyMask) & ~(uint32_t)(cpc - 1);
    const int s = cell_off[cb], m = cell_off[cb + cpc] - s;
    const int p = layout_rank(from, (int)(j - s), m);
    dst[s + layout_slot(to, p, m)] = v;
}

// Patch and cell key of loaded particles; all get direction 13 ("stay") so
// that the canonical order of a cell is the input order.
template <int ND>
__global__ void k_bin_initial(Geom g, PartPtrs A, int64_t n, uint32_t *__restrict__ keys)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x[3] = {A.x[i], A.y[i], ND == 3 ? A.z[i] : 0.};
    int pc[3] = {0, 0, 0}, ic[3] = {0, 0, 0};
    for (int d = 0; d < ND; ++d) {
        int c = (int)floor(x[d] / g.plen[d]);
        c = c < 0 ? 0 : (c > g.npatch[d] - 1 ? g.npatch[d] - 1 : c);
        while (c > 0 && x[d] < (double)c * g.plen[d]) --c;
        while (c < g.npatch[d] - 1 && x[d] >= (double)(c + 1) * g.plen[d]) ++c;
        pc[d] = c;
        int k = (int)floor(x[d] * (1.0 / g.dx[d])) - c * g.n[d];
        ic[d] = k < 0 ? 0 : (k > g.n[d] - 1 ? g.n[d] - 1 : k);
    }
    int key;
    if constexpr (ND == 3) {
        const int cl = ((ic[0] >> 2) * g.ncl[1] + (ic[1] >> 2)) * g.ncl[2] + (ic[2] >> 2);
        key = cl * 64 + 16 * (ic[0] & 3) + 4 * (ic[1] & 3) + (ic[2] & 3);
    } else {
        const int cl = (ic[0] >> 2) * g.ncl[1] + (ic[1] >> 2);
        key = cl * 16 + 4 * (ic[0] & 3) + (ic[1] & 3);
    }
    const int p = (pc[0] * g.npatch[1] + pc[1]) * g.npatch[2] + pc[2];
    keys[i] = (13u << 27) | (uint32_t)((int64_t)p * g.cpp + key);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int64_t sort_bsum_size(int64_t ncells) { return (ncells + kScanTile - 1) / kScanTile + 1; }

void launch_sort(const Geom &g, const uint32_t *keys, int64_t n, SortBufs &sb, hipStream_t st, bool counted,
                 bool in_cell_order, int layout, const int32_t *tie, bool placed)
{
    placed = placed && counted && !in_cell_order;
    const int64_t nc = g.ncells;
    if (n > 0 && !counted) {
        hipLaunchKernelGGL(k_hist, dim3(nblk(n, 256)), dim3(256), 0, st, keys, n, sb.counts);
        SMI_LAUNCH_CHECK();
    }
    const int64_t nb = (nc + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, st, (const int32_t *)sb.counts, nc,
                       sb.bsum);
    SMI_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(kScanThreads), 0, st, sb.bsum, nb);
    SMI_LAUNCH_CHECK();
    // the in-cell pass of the step's sort computes the ties of the particles
    // that stayed in their cluster from the clusters' previous ranges
    int stay13 = 1;
    for (int d = 0; d < g.nd; ++d)
        if (g.npatch[d] * g.n[d] <= 4) stay13 = 0;
    const TieSrc T{tie, counted && tie && in_cell_order ? sb.clus_old : nullptr, g.cells_per_cl,
                   (int)(nc / g.cells_per_cl), layout, n, stay13};
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, st, sb.counts, nc,
                       (const int32_t *)sb.bsum, sb.cell_off, placed ? sb.stay : nullptr,
                       const_cast<int32_t *>(T.clus_old), g.cells_per_cl);
    SMI_LAUNCH_CHECK();
    if (n > 0) {
        // cell-major places go straight to perm unless the canonical order
        // inside the cells and the interleaved layout are both asked for:
        // then they are ordered in tmp and laid out into perm
        const int cpc = layout ? g.cells_per_cl : 0;
        int32_t *first = in_cell_order && layout ? sb.tmp : sb.perm;
        hipLaunchKernelGGL(k_scatter, dim3(nblk(n, kScatTile)), dim3(kScatThreads), 0, st, keys, n, (const int32_t *)sb.cell_off,
                           sb.counts, first, in_cell_order ? 0 : cpc, layout, placed ? 1 : 0);
        SMI_LAUNCH_CHECK();
        if (in_cell_order) {
            if (sb.big) launch_fill_i32(sb.nbig, 2, 0, st);
            hipLaunchKernelGGL(k_bucket_fix, dim3(nblk(nc, 256)), dim3(256), 0, st, keys, nc,
                               (const int32_t *)sb.cell_off, first, sb.perm, cpc, layout, T, sb.big, sb.nbig);
            SMI_LAUNCH_CHECK();
            if (sb.big) {
                hipLaunchKernelGGL(k_bucket_fix_mid, dim3(nblk(nc, 256)), dim3(256), 0, st, keys, nc,
                                   (const int32_t *)sb.cell_off, first, sb.perm, cpc, layout, T, sb.big, sb.nbig);
                SMI_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_bucket_fix_big, dim3(kBigBlocks), dim3(kBigThreads), 0, st, keys,
                                   (const int32_t *)sb.cell_off, first, sb.perm, sb.tmp, cpc, layout, T,
                                   (const int32_t *)sb.big, sb.nbig);
                SMI_LAUNCH_CHECK();
            }
        }
    }
}

void launch_relayout(const Geom &g, const uint32_t *keys, int64_t n, SortBufs &sb, int from_layout,
                     int to_layout, hipStream_t st)
{
    if (n == 0 || from_layout == to_layout) return;
    hipLaunchKernelGGL(k_relayout, dim3(nblk(n, 256)), dim3(256), 0, st, keys, n, (const int32_t *)sb.cell_off,
                       (const int32_t *)sb.perm, sb.tmp, g.cells_per_cl, from_layout, to_layout);
    SMI_LAUNCH_CHECK();
    SMI_HIP(hipMemcpyAsync(sb.perm, sb.tmp, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
}

void launch_bin_initial(const Geom &g, const PartPtrs &A, int64_t n, uint32_t *keys, hipStream_t st)
{
    if (n == 0) return;
    if (g.nd == 3) hipLaunchKernelGGL(k_bin_initial<3>, dim3(nblk(n, 256)), dim3(256), 0, st, g, A, n, keys);
    else hipLaunchKernelGGL(k_bin_initial<2>, dim3(nblk(n, 256)), dim3(256), 0, st, g, A, n, keys);
    SMI_LAUNCH_CHECK();
}

}  // namespace smi
