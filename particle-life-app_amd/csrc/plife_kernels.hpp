// plife_kernels.hpp -- gfx950 kernels of the particle-life physics step.
//
// One step (reference: backend/Physics.java:112-134):
//   k_bin_tile               cell key per particle (Physics.java:362-375), cell histogram,
//                            first radix-digit histogram
//   k_scan2_*                exclusive/inclusive scans -> cell_start / cell_end
//                            (cell_end is the reference's `containers` array, Physics.java:329-348)
//   k_radix_hist/_scatter    stable LSD radix sort of (key, index): the reference's
//                            stable counting-sort permutation (Physics.java:342-353)
//   k_force (plife_force.hpp) velocity phase + position phase fused (Physics.java:397-450);
//                            legal because the update is Jacobi: velocities read only the
//                            sorted step-start positions, new positions go to another buffer.
//
// Positions are 32-bit fixed point (x = k * 2^-32, unit box), so the periodic
// minimum image of Range.wrapConnection (Range.java:68-81) is the exact int32
// difference k_q - k_p, and converting a position to FP64 is exact.  Cell keys
// use the same FP64 division and truncation as the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plife {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kRadixTile = 2048;                 // elements per radix tile
constexpr int kScatterBlock = 512;               // threads of the radix scatter
constexpr int kRadixItems = kRadixTile / kScatterBlock;
constexpr double kFix = 4294967296.0;       // 2^32
constexpr double kInvFix = 1.0 / 4294967296.0;

// ---------------------------------------------------------------------------
// wave64 / block helpers
// ---------------------------------------------------------------------------
// Inclusive scan over a wave64 in registers (DPP, no LDS round trips):
// Hillis-Steele steps of 1, 2, 4, 8 lanes inside each row of 16 (lanes
// shifted in from outside the row read 0), then row 0's total added to row 1
// and row 2's to row 3 (row_bcast:15), and the first two rows' total to rows
// 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    constexpr int kRowShr = 0x110, kBcast15 = 0x142, kBcast31 = 0x143;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 1, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 2, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 4, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr + 8, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kBcast15, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kBcast31, 0xc, 0xf, false);
    return v;
}

// Exclusive scan over a block of NT threads.  s_w: NT / 64 words of LDS.
template <int NT = kBlock>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w, uint32_t &total)
{
    const uint32_t inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) {
        uint32_t t = s_w[k];
        off += (k < w) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

// Sum over a block of NT threads; every thread gets the total.  s_w: NT / 64 words of LDS.
template <int NT = kBlock>
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v,
                                                            unsigned long long *s_w)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) s_w[w] = v;
    __syncthreads();
    unsigned long long tot = 0;
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) tot += s_w[k];
    __syncthreads();
    return tot;
}

// Block ids are dealt round-robin to the 8 XCDs; remap so each XCD works on a
// contiguous run of tiles and neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int q = nb >> 3, rem = nb & 7;
    const int xcd = b & 7, k = b >> 3;
    return xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
}

// ---------------------------------------------------------------------------
// fill
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint32_t v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The last step's counters to the host's pinned copy (the step after copies
// them itself, in k_bin_tile).
__global__ void k_copy_u64(const unsigned long long *__restrict__ src, unsigned long long *__restrict__ dst,
                           int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// cell keys (Physics.java:362-375): cx = trunc(x / cs) in FP64, clamped to nx-1.
// ---------------------------------------------------------------------------
// One world of a batch (a single-world handle is a batch of one).  Worlds
// occupy consecutive particle slots [off, off + n) in both the pre-sort and the
// sorted order, and consecutive cells [cell_base, cell_base + nx^DIM) of the
// batch's cell arrays, so one stable sort of the batch's cell keys sorts every
// world exactly as stepping it alone would.
struct WorldDev {
    int off, n;                      // particle slots
    int nx, all_global;              // cells per axis; 1 when nx < 3 (literal stencil)
    uint32_t cell_base;              // first batch cell of this world
    int ntypes, moff;                // matrix size and offset (doubles) in the batch matrix
    int wrap;
    float r2hi;                      // screen bound, units 2^-64: (rmax 2^32)^2 (1 + 2^-20)
    double cs;                       // cell size = rmax (Physics.java:312)
    double r2u;                      // fl(rmax*rmax) * 2^64: the reference's bound in units 2^-64
    double S;                        // 2^-32 / rmax: |rel| = |D| S with D in units 2^-32
    double R;                        // rmax * 2^32:  1/|rel| = R / |D|
    double beta, one_plus_beta, inv_1mb, inv_beta;
    double m2S;                      // -2 S
    double betaR;                    // beta R: |D| below it is the repulsion zone
    double ff;                       // friction^(60 dt) (host FP64 pow, Physics.java:401)
    double kdt;                      // force * dt * 2^-32
    double ka;                       // force * 2^-32: acceleration per unit of accumulated D (f/dist)
    double dt;
};

// World index of every slot of a batch: the last world whose first slot is
// at or before the slot (worlds are laid out in order; an empty world shares
// its offset with the next one, so the last such world is the non-empty one).
__global__ void k_world_of(int n, const WorldDev *__restrict__ worlds, int nworlds,
                           uint16_t *__restrict__ world_of)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nworlds;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (worlds[mid].off <= i) lo = mid;
        else hi = mid;
    }
    world_of[i] = (uint16_t)lo;
}

__device__ __forceinline__ int key_axis(uint32_t k, double cs, int nx)
{
    const double x = (double)k * kInvFix;   // exact
    int c = (int)(x / cs);                   // IEEE FP64 division, x >= 0: trunc == floor
    return c == nx ? nx - 1 : c;
}

// trunc(x / cs) for x = k 2^-32 without the division where that is safe:
// x * (1 / cs) is within 3 ulp of the rounded quotient, i.e. within 2^-35
// for quotients below 2^16 cells, so its integer part is the division's
// unless the product lies within 1e-9 of an integer -- then divide.
__device__ __forceinline__ int trunc_div_fast(uint32_t k, double cs, double ics)
{
    const double x = (double)k * kInvFix;
    const double q = x * ics;
    int c = (int)q;
    const double f = q - (double)c;
    if (f < 1e-9 || f > 1.0 - 1e-9) c = (int)(x / cs);
    return c;
}

__device__ __forceinline__ int key_axis_fast(uint32_t k, double cs, double ics, int nx)
{
    const int c = trunc_div_fast(k, cs, ics);
    return c == nx ? nx - 1 : c;
}

// Cell counts are accumulated with one atomic per run of equal keys inside a
// wave: the input is the previous step's cell order, so most lanes of a wave
// share a few keys and per-lane atomics would serialise on the same address.
__device__ __forceinline__ void count_runs(uint32_t key, bool valid, uint32_t *__restrict__ cell_count)
{
    const int lane = __lane_id();
    const uint32_t prev = __shfl_up(key, 1, kWave);
    const bool head = valid && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head);
    const int nvalid = (int)__popcll(__ballot(valid));   // valid lanes are a prefix
    if (head) {
        const uint64_t later = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
        const int next = later ? (int)__ffsll((long long)later) - 1 : nvalid;
        atomicAdd(&cell_count[key], (uint32_t)(next - lane));
    }
}

// Fused binning of one radix tile (kRadixTile elements in the radix kernels'
// element order): cell keys, the cell histogram (run-aggregated atomics) and
// this tile's histogram of the first radix digit.  Block 0 also clears the
// pair counter of the step.
template <int DIM, int D>
__global__ __launch_bounds__(1024) void k_bin_tile(const uint32_t *__restrict__ px,
                                                  const uint32_t *__restrict__ py,
                                                  const uint32_t *__restrict__ pz, int n,
                                                  const WorldDev *__restrict__ worlds,
                                                  const uint16_t *__restrict__ world_of,
                                                  int ntiles,
                                                  uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ cell_count,
                                                  uint32_t *__restrict__ hist,
                                                  unsigned long long *__restrict__ counters,
                                                  int ncounters, unsigned long long tag,
                                                  unsigned long long *__restrict__ stat_out)
{
    constexpr int R = 1 << D;
    __shared__ uint32_t h[R];
    for (int t = threadIdx.x; t < R; t += 1024) h[t] = 0;
    // the step's counters (pairs, sum of squares, per-world pairs) start at
    // zero, after the previous step's final values (and its layout tag) have
    // gone to the host's pinned copy (stat_out, when given)
    if (blockIdx.x == 0) {
        for (int t = threadIdx.x; t <= ncounters; t += 1024) {
            if (stat_out) stat_out[t] = counters[t];
            counters[t] = t < ncounters ? 0ull : tag;   // [ncounters]: which layout the counts belong to
        }
    }
    const double cs0 = worlds[0].cs;
    const int nx0 = worlds[0].nx;
    const uint32_t base0 = worlds[0].cell_base;
    __syncthreads();
    const int base = blockIdx.x * kRadixTile;
#pragma unroll
    for (int r = 0; r < kRadixTile / 1024; ++r) {
        const int i = base + r * 1024 + threadIdx.x;
        const bool valid = i < n;
        uint32_t key = 0xFFFFFFFFu;
        if (valid) {
            double cs = cs0;
            int nx = nx0;
            uint32_t cb = base0;
            if (world_of) {                            // several worlds: this slot's world
                const WorldDev *w = worlds + world_of[i];
                cs = w->cs;
                nx = w->nx;
                cb = w->cell_base;
            }
            key = (uint32_t)key_axis(px[i], cs, nx) + (uint32_t)key_axis(py[i], cs, nx) * (uint32_t)nx;
            if constexpr (DIM == 3) key += (uint32_t)key_axis(pz[i], cs, nx) * (uint32_t)nx * (uint32_t)nx;
            key += cb;
            keys[i] = key;
            atomicAdd(&h[key & (R - 1)], 1u);
        }
        count_runs(key, valid, cell_count);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R; t += 1024) hist[t * ntiles + blockIdx.x] = h[t];
}

// ---------------------------------------------------------------------------
// device-wide scan of u32 (reduce -> scan block sums -> downsweep)
// ---------------------------------------------------------------------------
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

// Two independent device-wide scans in the same three launches: blocks
// [0, nb0) scan job 0, blocks [nb0, nb0 + nb1) scan job 1.  excl may alias in;
// with zero_in (requires in != excl) the input is cleared once read, which
// leaves the cell counters ready for the next step's histogram.  With sq,
// the sum of squared inputs is added to *sq (for cell counts: the number of
// (particle, cell-mate) pairs, a clustering measure): the reduce pass stores
// each block's sum in sqb and the block-sum scan adds them up, so the
// device-wide total costs one atomic rather than one per block.
struct ScanJob {
    uint32_t *in;
    int n;
    uint32_t *excl;
    uint32_t *incl;   // may be null
    int zero_in;
    uint32_t *bsum;   // ceil(n / kScanTile) block sums
    unsigned long long *sq;   // may be null
    unsigned long long *sqb;  // ceil(n / kScanTile) per-block sums of squares (with sq)
};

__device__ __forceinline__ void scan_job_of(const ScanJob &j0, const ScanJob &j1, int nb0, int b,
                                            ScanJob &j, int &blk)
{
    if (b < nb0) { j = j0; blk = b; }
    else { j = j1; blk = b - nb0; }
}

__global__ __launch_bounds__(kBlock) void k_scan2_reduce(ScanJob j0, ScanJob j1, int nb0)
{
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_q[kBlock / kWave];
    ScanJob j;
    int blk;
    scan_job_of(j0, j1, nb0, blockIdx.x, j, blk);
    const int base = blk * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
    unsigned long long q = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < j.n) {
            const uint32_t v = j.in[base + k];
            s += v;
            q += (unsigned long long)v * v;
        }
    uint32_t total;
    block_excl_scan(s, s_w, total);
    if (threadIdx.x == 0) j.bsum[blk] = total;
    if (j.sq) {                                    // uniform per block
        q = block_sum_u64(q, s_q);
        if (threadIdx.x == 0) j.sqb[blk] = q;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan2_single(ScanJob j0, ScanJob j1)
{
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_q[kBlock / kWave];
    const ScanJob j = blockIdx.x == 0 ? j0 : j1;
    const int n = (j.n + kScanTile - 1) / kScanTile;
    uint32_t carry = 0;
    unsigned long long q = 0;
    for (int base = 0; base < n; base += kScanTile) {
        const int i0 = base + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (i0 + k < n) ? j.bsum[i0 + k] : 0u;
            s += v[k];
            if (j.sq && i0 + k < n) q += j.sqb[i0 + k];
        }
        uint32_t total;
        uint32_t off = block_excl_scan(s, s_w, total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < n) j.bsum[i0 + k] = off;
            off += v[k];
        }
        carry += total;
    }
    if (j.sq) {                                    // uniform per block
        q = block_sum_u64(q, s_q);
        if (threadIdx.x == 0 && q) atomicAdd(j.sq, q);
    }
}

__global__ __launch_bounds__(kBlock) void k_scan2_down(ScanJob j0, ScanJob j1, int nb0)
{
    __shared__ uint32_t s_w[4];
    ScanJob j;
    int blk;
    scan_job_of(j0, j1, nb0, blockIdx.x, j, blk);
    const int i0 = blk * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (i0 + k < j.n) ? j.in[i0 + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t off = block_excl_scan(s, s_w, total) + j.bsum[blk];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (i0 + k < j.n) {
            if (j.zero_in) j.in[i0 + k] = 0u;
            j.excl[i0 + k] = off;
            off += v[k];
            if (j.incl) j.incl[i0 + k] = off;
        }
    }
}

// Down-sweep for jobs of at most kFoldBlocks blocks, with the block-sum scan
// folded in: each block adds up the sums of the blocks before it (at most
// kFoldBlocks / kBlock loads per thread), and block 0 adds the job's sum of
// squares -- two launches instead of three, same outputs.
constexpr int kFoldBlocks = 1024;

__global__ __launch_bounds__(kBlock) void k_scan2_down_fold(ScanJob j0, ScanJob j1, int nb0)
{
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_q[kBlock / kWave];
    ScanJob j;
    int blk;
    scan_job_of(j0, j1, nb0, blockIdx.x, j, blk);
    uint32_t pre = 0;
    for (int b = threadIdx.x; b < blk; b += kBlock) pre += j.bsum[b];
    uint32_t before;
    block_excl_scan(pre, s_w, before);             // `before` = the sum over all threads
    const int i0 = blk * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (i0 + k < j.n) ? j.in[i0 + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t off = block_excl_scan(s, s_w, total) + before;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (i0 + k < j.n) {
            if (j.zero_in) j.in[i0 + k] = 0u;
            j.excl[i0 + k] = off;
            off += v[k];
            if (j.incl) j.incl[i0 + k] = off;
        }
    }
    if (j.sq && blk == 0) {                        // uniform per block
        const int nb = (j.n + kScanTile - 1) / kScanTile;
        unsigned long long q = 0;
        for (int b = threadIdx.x; b < nb; b += kBlock) q += j.sqb[b];
        q = block_sum_u64(q, s_q);
        if (threadIdx.x == 0 && q) atomicAdd(j.sq, q);
    }
}

// Both jobs in one launch when they are small (at most kSmallScan entries):
// workgroup j scans job j alone, kScanItems entries per thread per pass,
// carrying the running sum -- one launch instead of three, for arrays whose
// scan time is launch latency.  Same outputs as the three-launch scan.
constexpr int kSmallBlock = 1024;
constexpr int kSmallScan = 8192;            // at most two passes of the workgroup

__global__ __launch_bounds__(kSmallBlock) void k_scan_small(ScanJob j0, ScanJob j1)
{
    __shared__ uint32_t s_w[kSmallBlock / kWave];
    __shared__ unsigned long long s_q[kSmallBlock / kWave];
    const ScanJob j = blockIdx.x == 0 ? j0 : j1;
    uint32_t carry = 0;
    unsigned long long q = 0;
    for (int base = 0; base < j.n; base += kSmallBlock * kScanItems) {
        const int i0 = base + threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (i0 + k < j.n) ? j.in[i0 + k] : 0u;
            s += v[k];
            q += (unsigned long long)v[k] * v[k];
        }
        uint32_t total;
        uint32_t off = block_excl_scan<kSmallBlock>(s, s_w, total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < j.n) {
                if (j.zero_in) j.in[i0 + k] = 0u;
                j.excl[i0 + k] = off;
                off += v[k];
                if (j.incl) j.incl[i0 + k] = off;
            }
        }
        carry += total;
    }
    if (j.sq) {                                    // uniform per block
        q = block_sum_u64<kSmallBlock>(q, s_q);
        if (threadIdx.x == 0 && q) atomicAdd(j.sq, q);
    }
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (key, value) -- reduce-then-scan per digit pass
// ---------------------------------------------------------------------------

template <int D>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t *__restrict__ keys, int n,
                                                       int shift, int ntiles,
                                                       uint32_t *__restrict__ hist)
{
    constexpr int R = 1 << D;
    __shared__ uint32_t h[R];
    for (int t = threadIdx.x; t < R; t += kBlock) h[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRadixTile;
#pragma unroll 4
    for (int k = 0; k < kRadixTile / kBlock; ++k) {
        const int i = base + k * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (R - 1)], 1u);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R; t += kBlock) hist[t * ntiles + blockIdx.x] = h[t];
}

// ---------------------------------------------------------------------------
// SoA state
// ---------------------------------------------------------------------------
// Compaction of a batch after worlds left: destination slot i of world w
// (worlds[w].off <= i < worlds[w].off + worlds[w].n in the new layout) takes
// the particle at src_off[w] + (i - worlds[w].off) of the old layout.
__global__ __launch_bounds__(kBlock) void k_compact(int n, const WorldDev *__restrict__ worlds,
                                                   const uint16_t *__restrict__ world_of,
                                                   const int *__restrict__ src_off,
                                                   const uint32_t *__restrict__ sx,
                                                   const uint32_t *__restrict__ sy,
                                                   const uint32_t *__restrict__ sz,
                                                   const float *__restrict__ svx,
                                                   const float *__restrict__ svy,
                                                   const float *__restrict__ svz,
                                                   const int32_t *__restrict__ stype,
                                                   const int32_t *__restrict__ sid,
                                                   uint32_t *__restrict__ dx, uint32_t *__restrict__ dy,
                                                   uint32_t *__restrict__ dz, float *__restrict__ dvx,
                                                   float *__restrict__ dvy, float *__restrict__ dvz,
                                                   int32_t *__restrict__ dtype, int32_t *__restrict__ did)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int w = world_of ? (int)world_of[i] : 0;
    const int j = src_off[w] + (i - worlds[w].off);
    dx[i] = sx[j];
    dy[i] = sy[j];
    dz[i] = sz[j];
    dvx[i] = svx[j];
    dvy[i] = svy[j];
    dvz[i] = svz[j];
    dtype[i] = stype[j];
    did[i] = sid[j];
}

// Seeded world generation, the device side of the default presets
// (backend/DefaultPositionSetter.java: uniform positions; backend/
// DefaultTypeSetter.java: uniform types; velocities 0, Physics.java:297-303).
// Counter-based SplitMix64 (Vigna's published generator): draw k of stream s
// of `seed` is mix(base_s + (k + 1) * gamma), base_s = mix(seed + s * kStream).
// Uniforms are the top 24 bits (a 2^-24 grid, exact in FP32, FP64 and the
// 32-bit fixed point); streams 1..3 are x, y, z and 4 the type.
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kStream = 0xD1B54A32D192ED03ull;

__host__ __device__ inline uint64_t splitmix_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct GenArgs {
    uint64_t base[4];                  // stream bases for x, y, z, type
    int n, off, ntypes, dim;
    uint32_t *px, *py, *pz;
    float *vx, *vy, *vz;
    int32_t *type, *id;
};

__global__ __launch_bounds__(kBlock) void k_gen_uniform(GenArgs g)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= g.n) return;
    const uint64_t step = (uint64_t)(k + 1) * kGamma;
    const int i = g.off + k;
    g.px[i] = (uint32_t)(splitmix_mix(g.base[0] + step) >> 40) << 8;      // u24 * 2^-24 in 2^-32 units
    g.py[i] = (uint32_t)(splitmix_mix(g.base[1] + step) >> 40) << 8;
    g.pz[i] = g.dim == 3 ? (uint32_t)(splitmix_mix(g.base[2] + step) >> 40) << 8 : 0u;
    const uint64_t m = splitmix_mix(g.base[3] + step) >> 40;            // floor(u24 * T) exactly
    g.type[i] = (int32_t)((m * (uint64_t)g.ntypes) >> 24);
    g.vx[i] = 0.0f;
    g.vy[i] = 0.0f;
    g.vz[i] = 0.0f;
    g.id[i] = k;
}

// Physics.ensurePosition (Physics.java:499-505) of an FP64 coordinate, then
// rounding to the 2^-32 grid.  Periodic: the fractional part (a result of
// 1.0, reference quirk Q3, is the same torus point as 0).  Clamped: [0, 1],
// with 1.0 stored as 1 - 2^-32.  NaN becomes 0.
__host__ __device__ inline uint32_t fixed_of(double x, bool wrap)
{
    if (!(x == x)) return 0u;
    if (wrap) {
        x = x - floor(x);
        return (uint32_t)(unsigned long long)rint(x * kFix);
    }
    x = x < 0 ? 0 : (x > 1 ? 1 : x);
    const double s = rint(x * kFix);
    return s >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)(unsigned long long)s;
}

// ---------------------------------------------------------------------------
// The app's position and type presets (app/PositionSetterProvider.java:17-119,
// app/TypeSetterProvider.java:12-45) over a range of slots, with the
// counter-based SplitMix64 draws above in place of the reference's unseeded
// java.util.Random: uniform u = (draw >> 11) * 2^-53, Gaussians by Box-Muller
// from two streams, sqrt(-2 ln(1 - u1)) cos(2 pi u2).  Shapes are computed in
// the reference's side-2 frame with its float constants and evaluation order,
// mapped to the unit box by x * 0.5 + 0.5 as each preset does, then placed
// by fixed_of.  The uniform preset is the generator above (2^-24 grid,
// streams 1-3).  3D worlds (no reference counterpart) take z from stream 3.
// ---------------------------------------------------------------------------
enum PositionPreset {
    kPosCentered, kPosUniform, kPosUniformCircle, kPosCenteredCircle, kPosRing, kPosRainbowRing,
    kPosColorBattle, kPosColorWheel, kPosLine, kPosSpiral, kPosRainbowSpiral, kPosPresets
};
enum TypePreset {
    kTypeRandom, kTypeRandomize10, kTypeSlices, kTypeOnion, kTypeRotate, kTypeFlip,
    kTypeMoreOfFirst, kTypeKillStill, kTypePresets
};
constexpr int kPresetStream = 40;      // preset draws use streams 40..43 (positions), 44..45 (types)
constexpr double kPi = 3.141592653589793;     // java.lang.Math.PI

struct PresetArgs {
    uint64_t base[6];                  // streams 40..45
    uint64_t ubase[4];                 // streams 1..4 (the uniform generator's)
    int n, off, ntypes, dim, preset, wrap;
    uint32_t *px, *py, *pz;
    float *vx, *vy, *vz;
    int32_t *type;
};

__device__ __forceinline__ double preset_u(uint64_t base, uint64_t step)
{
    return (double)(splitmix_mix(base + step) >> 11) * 0x1p-53;
}

__device__ __forceinline__ double preset_gauss(uint64_t b1, uint64_t b2, uint64_t step)
{
#pragma clang fp contract(off)
    const double u1 = ((double)(splitmix_mix(b1 + step) >> 11) + 1.0) * 0x1p-53;   // (0, 1]
    const double u2 = (double)(splitmix_mix(b2 + step) >> 11) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
}

// mapType of TypeSetterProvider.java:46-52
__device__ __forceinline__ int preset_map_type(double v, int nt)
{
    return max(0, min(nt - 1, (int)floor(v * nt)));
}

__global__ __launch_bounds__(kBlock) void k_set_positions(PresetArgs g)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= g.n) return;
    const int i = g.off + k;
    const uint64_t s = (uint64_t)(k + 1) * kGamma;
    const int t = g.type[i];
    const double T = (double)g.ntypes;
    double x = 0.0, y = 0.0;
    switch (g.preset) {
        case kPosCentered: {
            const double scale = (double)0.3f;
            x = preset_gauss(g.base[0], g.base[1], s) * scale;
            y = preset_gauss(g.base[2], g.base[3], s) * scale;
            break;
        }
        case kPosUniformCircle: {
            const double angle = preset_u(g.base[0], s) * 2 * kPi;
            const double radius = 0.5 * sqrt(preset_u(g.base[1], s));
            x = cos(angle) * radius;
            y = sin(angle) * radius;
            break;
        }
        case kPosCenteredCircle: {
            const double angle = preset_u(g.base[0], s) * 2 * kPi;
            const double radius = (double)0.5f * preset_u(g.base[1], s);
            x = cos(angle) * radius;
            y = sin(angle) * radius;
            break;
        }
        case kPosRing: {
            const double angle = preset_u(g.base[0], s) * 2 * kPi;
            const double radius = 0.7 + 0.02 * preset_gauss(g.base[2], g.base[3], s);
            x = cos(angle) * radius;
            y = sin(angle) * radius;
            break;
        }
        case kPosRainbowRing: {
            const double angle = (0.3 * preset_gauss(g.base[0], g.base[1], s) + t) / T * 2 * kPi;
            const double radius = 0.7 + 0.02 * preset_gauss(g.base[2], g.base[3], s);
            x = cos(angle) * radius;
            y = sin(angle) * radius;
            break;
        }
        case kPosColorBattle: {
            const double ca = t / T * 2 * kPi;
            const double cr = (double)0.5f;
            const double angle = preset_u(g.base[0], s) * 2 * kPi;
            const double radius = preset_u(g.base[1], s) * (double)0.1f;
            x = cr * cos(ca) + cos(angle) * radius;
            y = cr * sin(ca) + sin(angle) * radius;
            break;
        }
        case kPosColorWheel: {
            const double ca = t / T * 2 * kPi;
            const double cr = (double)0.3f, ir = (double)0.2f;
            x = cr * cos(ca) + preset_gauss(g.base[0], g.base[1], s) * ir;
            y = cr * sin(ca) + preset_gauss(g.base[2], g.base[3], s) * ir;
            break;
        }
        case kPosLine:
            x = 2 * preset_u(g.base[0], s) - 1;
            y = (2 * preset_u(g.base[1], s) - 1) * (double)0.15f;
            break;
        case kPosSpiral:
        case kPosRainbowSpiral: {
            double f;
            if (g.preset == kPosSpiral) {
                f = preset_u(g.base[0], s);
            } else {
                const double spread_t = 0.3 / g.ntypes;
                f = (t + 1) / (double)(g.ntypes + 2) + spread_t * preset_gauss(g.base[0], g.base[1], s);
                f = f < 0 ? 0 : (f > 1 ? 1 : f);
            }
            const double angle = 2.0 * 2 * kPi * f;
            const double spread = 0.5 * fmin(f, 0.2);
            const double radius = 0.9 * f + spread * preset_gauss(g.base[2], g.base[3], s) * spread;
            x = radius * cos(angle);
            y = radius * sin(angle);
            break;
        }
        default:    // kPosUniform: the uniform generator's grid values
            x = (double)(splitmix_mix(g.ubase[0] + s) >> 40) * 0x1p-24;
            y = (double)(splitmix_mix(g.ubase[1] + s) >> 40) * 0x1p-24;
            break;
    }
    if (g.preset != kPosUniform) {
        x = x * 0.5 + 0.5;
        y = y * 0.5 + 0.5;
    }
    const bool wrap = g.wrap != 0;
    g.px[i] = fixed_of(x, wrap);
    g.py[i] = fixed_of(y, wrap);
    g.pz[i] = g.dim == 3 ? (uint32_t)(splitmix_mix(g.ubase[2] + s) >> 40) << 8 : 0u;
    g.vx[i] = 0.0f;     // Physics.setPosition (Physics.java:297-303)
    g.vy[i] = 0.0f;
    g.vz[i] = 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_set_types(PresetArgs g)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= g.n) return;
    const int i = g.off + k;
    const uint64_t s = (uint64_t)(k + 1) * kGamma;
    const int nt = g.ntypes;
    int t = g.type[i];
    switch (g.preset) {
        case kTypeRandomize10:
            if (preset_u(g.base[4], s) < 0.1) t = preset_map_type(preset_u(g.base[5], s), nt);
            break;
        case kTypeSlices:
            t = preset_map_type((double)g.px[i] * kInvFix, nt);
            break;
        case kTypeOnion: {
            const double dx = (double)g.px[i] * kInvFix - 0.5, dy = (double)g.py[i] * kInvFix - 0.5;
            const double dz = g.dim == 3 ? (double)g.pz[i] * kInvFix : 0.0;
            t = preset_map_type(sqrt(dx * dx + (dy * dy + dz * dz)) * 2, nt);
            break;
        }
        case kTypeRotate:
            t = (t + 1) % nt;
            break;
        case kTypeFlip:
            t = nt - 1 - t;
            break;
        case kTypeMoreOfFirst:
            t = preset_map_type(preset_u(g.base[4], s) * preset_u(g.base[5], s), nt);
            break;
        case kTypeKillStill: {
            const double vx = g.vx[i], vy = g.vy[i], vz = g.dim == 3 ? (double)g.vz[i] : 0.0;
            if (sqrt(vx * vx + (vy * vy + vz * vz)) < 0.01) t = nt - 1;
            break;
        }
        default: {  // kTypeRandom: the uniform generator's type draw
            const uint64_t m = splitmix_mix(g.ubase[3] + s) >> 40;
            t = (int32_t)((m * (uint64_t)nt) >> 24);
            break;
        }
    }
    g.type[i] = t;
}

// One in-place edit of a particle between steps (positions already wrapped or
// clamped and rounded to the 2^-32 grid); ids are kept.
struct SlotEdit {
    int32_t slot;
    uint32_t x, y, z;
    float vx, vy, vz;
    int32_t type;
};

__global__ __launch_bounds__(kBlock) void k_set_slots(const SlotEdit *__restrict__ e, int k, int off,
                                                      uint32_t *px, uint32_t *py, uint32_t *pz,
                                                      float *vx, float *vy, float *vz, int32_t *type)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= k) return;
    const SlotEdit r = e[i];
    const int s = off + r.slot;
    px[s] = r.x;
    py[s] = r.y;
    pz[s] = r.z;
    vx[s] = r.vx;
    vy[s] = r.vy;
    vz[s] = r.vz;
    type[s] = r.type;
}

struct PosBuf {
    uint32_t *x, *y, *z;
};
struct AuxBuf {
    float *vx, *vy, *vz;
    int32_t *type, *id;
};

// Stable scatter.  Elements are ranked in index order: round r covers indices
// base + r*256 + tid, and inside a round waves go in order and lanes go in
// order (wave-level match by digit ballots).  gofs is the digit-major
// exclusive scan of the per-tile histograms.  vin == nullptr means value = index.
template <int D>
__global__ __launch_bounds__(kScatterBlock) void k_radix_scatter(const uint32_t *__restrict__ kin,
                                                                 const uint32_t *__restrict__ vin,
                                                                 uint32_t *__restrict__ kout,
                                                                 uint32_t *__restrict__ vout,
                                                                 const uint32_t *__restrict__ gofs,
                                                                 int n, int shift, int ntiles)
{
    constexpr int R = 1 << D;
    constexpr int NW = kScatterBlock / kWave;
    __shared__ uint32_t s_k[kRadixTile];
    __shared__ uint32_t s_v[kRadixTile];
    __shared__ uint32_t s_cnt[NW][R];
    __shared__ uint32_t s_off[NW][R];
    __shared__ uint32_t s_start[R];
    __shared__ uint32_t s_gb[R];
    __shared__ uint32_t s_w[NW];

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int base = blockIdx.x * kRadixTile;
    for (int t = tid; t < NW * R; t += kScatterBlock) (&s_cnt[0][0])[t] = 0;
    uint32_t running = 0;
    uint32_t key[kRadixItems], val[kRadixItems], rank[kRadixItems];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const int i = base + r * kScatterBlock + tid;
        const bool valid = i < n;
        const uint32_t k = valid ? kin[i] : 0u;
        const uint32_t v = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
        const uint32_t d = (k >> shift) & (R - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < D; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t lr = (uint32_t)__popcll(peers & lt);
        if (valid && lr == 0) s_cnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (tid < R) {
            uint32_t o = running;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) {
                s_off[ww][tid] = o;
                o += s_cnt[ww][tid];
                s_cnt[ww][tid] = 0;
            }
            running = o;
        }
        __syncthreads();
        key[r] = k;
        val[r] = v;
        rank[r] = valid ? s_off[w][d] + lr : 0u;
    }

    // exclusive scan of the tile's digit counts (threads tid < R hold them)
    const uint32_t rv = tid < R ? running : 0u;
    const uint32_t rinc = wave_incl_scan(rv);
    if (lane == 63) s_w[w] = rinc;
    __syncthreads();
    uint32_t wb = 0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) wb += ww < w ? s_w[ww] : 0u;
    const uint32_t st = wb + rinc - rv;
    if (tid < R) {
        s_start[tid] = st;
        s_gb[tid] = gofs[tid * ntiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const int i = base + r * kScatterBlock + tid;
        if (i < n) {
            const uint32_t d = (key[r] >> shift) & (R - 1);
            const uint32_t lp = s_start[d] + rank[r];
            s_k[lp] = key[r];
            s_v[lp] = val[r];
        }
    }
    __syncthreads();
    const int nvalid = min(kRadixTile, n - base);
    for (int j = tid; j < nvalid; j += kScatterBlock) {
        const uint32_t k = s_k[j];
        const uint32_t d = (k >> shift) & (R - 1);
        const uint32_t g = s_gb[d] + (uint32_t)j - s_start[d];
        if (kout) kout[g] = k;
        vout[g] = s_v[j];
    }
}

// Slots [off, off + n) of up to kPackMax device arrays of 32-bit words,
// gathered chunk-major for a chunked download: chunk c (slots c*chunk ...)
// holds each array's len_c words in turn, so one DMA moves one chunk.
constexpr int kPackMax = 8;
struct PackArgs {
    const uint32_t *src[kPackMax];
    int na, off, n, chunk;
};

__global__ __launch_bounds__(kBlock) void k_pack(PackArgs p, uint32_t *dst)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= p.n) return;
    const int c = i / p.chunk, j = i - c * p.chunk;
    const int len = min(p.chunk, p.n - c * p.chunk);
    uint32_t *d = dst + (size_t)c * p.chunk * p.na + j;
#pragma unroll
    for (int a = 0; a < kPackMax; ++a)
        if (a < p.na) d[(size_t)a * len] = p.src[a][p.off + i];
}

// The inverse for uploads: staged arrays x, y, [z], vx, vy, [vz], type (each
// n words, FP32 velocities as their bits) into slots [off, off + n) of the
// SoA buffers, ids 0..n-1, z and vz zero in 2D.
struct UnpackArgs {
    uint32_t *px, *py, *pz;
    float *vx, *vy, *vz;
    int32_t *type, *id;
    int off, n, dim;
};

__global__ __launch_bounds__(kBlock) void k_unpack(UnpackArgs u, const uint32_t *__restrict__ src)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= u.n) return;
    const size_t n = (size_t)u.n;
    const bool d3 = u.dim == 3;
    const uint32_t *v = src + (d3 ? 3 : 2) * n;      // velocities
    const int s = u.off + i;
    u.px[s] = src[i];
    u.py[s] = src[n + i];
    u.pz[s] = d3 ? src[2 * n + i] : 0u;
    u.vx[s] = __uint_as_float(v[i]);
    u.vy[s] = __uint_as_float(v[n + i]);
    u.vz[s] = d3 ? __uint_as_float(v[2 * n + i]) : 0.0f;
    u.type[s] = (int32_t)v[(d3 ? 3 : 2) * n + i];
    u.id[s] = i;
}

}  // namespace plife
