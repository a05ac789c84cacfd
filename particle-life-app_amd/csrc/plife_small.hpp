// plife_small.hpp -- the one-launch step of worlds that fit one workgroup's LDS.
//
// Reference: backend/Physics.java:112-134 (update), 309-375 (makeContainers,
// getContainerIndex), 397-450 (updateVelocity, updatePosition).
//
// A world of up to about ten thousand particles (2D; fewer in 3D) fits the
// 160 KB of LDS a gfx950 workgroup may hold: its cell keys, the stable sort
// of (key, index), the cell table and the sorted positions and types.  One
// launch then steps every world of the handle with no binning, scan or sort
// launches: a world of n particles takes ceil(n / homes) workgroups of 1024
// threads.  Each workgroup bins and sorts the whole world itself (a few
// microseconds of one CU; sharing the sort would need a grid-wide barrier)
// and steps `homes` consecutive particles of the sorted order: with at most
// 256 homes, 4 or 16 lanes share each home (one lane's chain of dependent
// loads and FP64 steps is the step's critical path), else 1024 homes at a
// time; with many worlds a workgroup owns a whole world.  The sort is
// incremental where it can be: the slots are the previous step's sorted
// order, so only the particles that changed cell are ranked.  Every home
// visits its candidates in the order of the other force kernels (home_global:
// per stencil row, cells left to right, each cell in sorted order) with the
// same FP32 superset screen, FP64 force law and integrator, so the new state
// is bit-identical to the multi-launch path's (the lanes that share a home
// add its terms in the same order; the FP32 screen only skips pairs the FP64
// predicate rejects).
//
// A step is a chain of short phases on one CU, so per-thread loops are
// unrolled with clamped indices: the passes over the particles issue eight
// loads per thread together, the other loops four, instead of one round trip
// per iteration.  Loops are not unrolled further, and the boundary mode is a
// run-time flag, to keep the code (run once per workgroup) small.
//
// Counters: workgroups add their pairs into an accumulator; the last
// workgroup to finish (a ticket counter, no waiting) publishes the step's
// counters and the host's pinned copy and clears the accumulator for the
// next step.
#pragma once

#include "plife_force.hpp"

namespace plife {

constexpr int kSmallThreads = 1024;
constexpr int kSmallWaves = kSmallThreads / kWave;
constexpr int kSmallDigitBits = 7;           // radix digit of the in-LDS sort
constexpr int kSmallPer = 16;                // particles per thread at most
constexpr int kSmallU = 4;                   // loop unroll: loads issued together
constexpr int kSmallL = 8;                   // the same for the passes over the particles
constexpr int kSmallMaxN = kSmallThreads * kSmallPer;   // also below 2^16: entries pack (key << 16 | index)
constexpr int kSmallMaxCells = 1 << 14;      // keys of at most 14 bits: two radix passes

struct SmallArgs {
    ForceArgs f;                     // state, outputs, matrix, worlds; blocks = (world, first home slot)
    uint32_t *perm_out;              // sorted slot -> pre-sort index
    const uint32_t *skeys_in;        // the previous step's skeys_out (a hint: checked before use)
    uint32_t *skeys_out;             // batch cell key of each sorted slot
    uint32_t *keys_out;              // batch cell key of each pre-sort index
    uint32_t *cell_start_out, *cell_end_out;
    unsigned long long *acc;         // [0] pairs, [1] sum of count^2, [2 + w] pairs of world w; [ncounters] tickets
    unsigned long long *counters;    // the step's counters, [ncounters] = tag
    unsigned long long *stat_out;    // the host's pinned copy (may be null)
    unsigned long long tag;
    int ncounters;
    int homes;                       // homes per workgroup (any count; <= 64: 16 lanes per home, <= 256: 4)
    int n_max, cells_max;            // largest world of the launch (LDS layout)
};

struct SmallLds {
    size_t m, p0, p1, cnt, cst, sy, sz, st, total;
};

// The sorted entries ping-pong between p0 and p1; after the sort the buffer
// that does not hold the result is reused for the sorted x coordinates.  The
// incremental sort keeps its scratch (stayer prefix, movers) in the sy/sz/st
// region, which the gather fills only afterwards.
__host__ __device__ inline SmallLds small_lds_layout(int dim, int ntypes_lds, int n_max, int cells_max)
{
    SmallLds L;
    size_t o = 0;
    L.m = o;    o += (size_t)ntypes_lds * ntypes_lds * 8;
    L.p0 = o;   o += (size_t)n_max * 4;
    L.p1 = o;   o += (size_t)n_max * 4;
    L.cnt = o;  o += (size_t)kSmallWaves * (1 << kSmallDigitBits) * 4;
    L.cst = o;  o += (size_t)(cells_max + 1) * 2;
    o = (o + 3) & ~(size_t)3;
    L.sy = o;   o += (size_t)n_max * 4;
    L.sz = o;   o += dim == 3 ? (size_t)n_max * 4 : 0;
    L.st = o;   o += (size_t)n_max;
    L.total = (o + 15) & ~(size_t)15;
    return L;
}

// Candidates lo..hi-1 of the world's sorted order for the home at sorted
// position hl: visit_run over LDS in two phases per 32 candidates.  The FP32
// screen of k_force (exact int32 deltas, |D|^2 <= rmax^2 (1 + 2^-20), a strict
// superset of the FP64 predicate) runs kSmallU candidates at a time with their
// loads issued together and sets one bit per kept candidate; then the kept
// candidates, in order, take the reference's FP64 predicate, the clamped
// world's sign test and the force law.  The sums are those of visit_run.
template <int DIM, bool DEBUG, bool MLDS>
__device__ __forceinline__ void small_run(const ForceArgs &a, const WorldDev &w, bool wrap, const double *M,
                                          const uint32_t *SX, const uint32_t *SY, const uint32_t *SZ,
                                          const uint8_t *ST, const uint32_t *P, int hl, uint32_t hx,
                                          uint32_t hy, uint32_t hz, int row, Acc &acc, int lo, int hi)
{
#pragma unroll 1
    for (int c0 = lo; c0 < hi; c0 += 32) {
        const int cn = min(32, hi - c0);
        uint32_t mask = 0;
#pragma unroll 1
        for (int u0 = 0; u0 < cn; u0 += kSmallU) {
            uint32_t qx[kSmallU], qy[kSmallU], qz[kSmallU];
#pragma unroll
            for (int u = 0; u < kSmallU; ++u) {
                const int j = min(c0 + u0 + u, hi - 1);
                qx[u] = SX[j];
                qy[u] = SY[j];
                qz[u] = DIM == 3 ? SZ[j] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kSmallU; ++u)
                if (u0 + u < cn && screen<DIM>(hx, hy, hz, qx[u], qy[u], qz[u], w.r2hi)) mask |= 1u << (u0 + u);
        }
        if (hl >= c0 && hl < c0 + 32) mask &= ~(1u << (hl - c0));
        while (mask) {
            const int j = c0 + __builtin_ctz(mask);
            mask &= mask - 1u;
            const uint32_t qx = SX[j], qy = SY[j], qz = DIM == 3 ? SZ[j] : 0u;
            const int dxi = (int)(qx - hx), dyi = (int)(qy - hy);
            const int dzi = DIM == 3 ? (int)(qz - hz) : 0;
            if (!wrap) {
                // plain difference of a clamped world: skip |D| >= 1/2
                bool ok = ((dxi >= 0) == (qx >= hx)) & ((dyi >= 0) == (qy >= hy));
                if (DIM == 3) ok &= ((dzi >= 0) == (qz >= hz));
                if (!ok) continue;
            }
            const int idx = row + ST[j];
            const double m = MLDS ? M[idx] : a.matrix[w.moff + idx] * w.R;
            pair64<DIM, DEBUG>(acc, dxi, dyi, dzi, m,
                               DEBUG ? (uint32_t)a.id[w.off + min((int)(P[j] & 0xFFFFu), w.n - 1)] : 0u, w);
        }
    }
}

// The same candidates with G lanes per home (G = 4: a quad of lanes; G = 16:
// a row of 16), for workgroups with fewer homes than threads: lane sub of the
// home's group evaluates candidate j0 + sub of each step of G candidates --
// the reference's FP64 predicate, the clamped world's sign test and the force
// law, with no screen and no branch (c = 0 where the pair does not interact)
// -- and then every lane of the group adds the step's G terms in candidate
// order, acc = fma(c_k, D_k, acc) for k = 0 .. G-1, with each lane's c and D
// broadcast across the group by DPP.  These are small_run's operations in its
// order (a term with c = 0 leaves a sum unchanged: a sum that starts at +0 is
// never -0), so the sums are bit-identical, while one home's chain of
// dependent loads and FP64 steps is shared by G lanes.  Interacting pairs
// (and the debug hash) are counted per lane; the caller sums the group.
template <int G, int K>
__device__ __forceinline__ uint32_t group_bcast(uint32_t v)
{
    static_assert(G == 4 || G == 16, "groups are DPP quads or rows");
    if constexpr (G == 4) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, K * 0x55, 0xf, 0xf, false);  // quad_perm:[K,K,K,K]
    else return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x150 + K, 0xf, 0xf, false);  // row_newbcast:K
}

template <int DIM, int G, int K = 0>
__device__ __forceinline__ void group_terms(Acc &acc, double c, int dx, int dy, int dz)
{
    if constexpr (K < G) {
        const unsigned long long cb = (unsigned long long)__double_as_longlong(c);
        const unsigned long long kb = (unsigned long long)group_bcast<G, K>((uint32_t)cb) |
                                      ((unsigned long long)group_bcast<G, K>((uint32_t)(cb >> 32)) << 32);
        const double ck = __longlong_as_double((long long)kb);
        acc.x = fma(ck, (double)(int)group_bcast<G, K>((uint32_t)dx), acc.x);
        acc.y = fma(ck, (double)(int)group_bcast<G, K>((uint32_t)dy), acc.y);
        if constexpr (DIM == 3) acc.z = fma(ck, (double)(int)group_bcast<G, K>((uint32_t)dz), acc.z);
        group_terms<DIM, G, K + 1>(acc, c, dx, dy, dz);
    }
}

template <int DIM, bool DEBUG, bool MLDS, int G>
__device__ __forceinline__ void small_run_split(const ForceArgs &a, const WorldDev &w, bool wrap,
                                                const double *M, const uint32_t *SX, const uint32_t *SY,
                                                const uint32_t *SZ, const uint8_t *ST, const uint32_t *P,
                                                int sub, uint32_t hx, uint32_t hy, uint32_t hz, int row,
                                                Acc &acc, int lo, int hi)
{
#pragma unroll 1
    for (int j0 = lo; j0 < hi; j0 += G) {
        const int j = j0 + sub;
        const int jc = min(j, hi - 1);
        const uint32_t qx = SX[jc], qy = SY[jc], qz = DIM == 3 ? SZ[jc] : 0u;
        const int idx = row + ST[jc];
        const double m = MLDS ? M[idx] : a.matrix[w.moff + idx] * w.R;
        const int dxi = (int)(qx - hx), dyi = (int)(qy - hy);
        const int dzi = DIM == 3 ? (int)(qz - hz) : 0;
        bool in = j < hi;
        if (!wrap) {
            // plain difference of a clamped world: skip |D| >= 1/2
            in &= ((dxi >= 0) == (qx >= hx)) & ((dyi >= 0) == (qy >= hy));
            if (DIM == 3) in &= ((dzi >= 0) == (qz >= hz));
        }
        const double d2 = pair_d2<DIM>((double)dxi, (double)dyi, (double)dzi);
        in &= (d2 != 0.0) & (d2 <= w.r2u);           // Physics.java:432
        const double c = in ? pair_coef(d2, m, w) : 0.0;
        acc.cnt += in ? 1 : 0;
        if constexpr (DEBUG)
            if (in) acc.hash += id_hash(a.id[w.off + min((int)(P[jc] & 0xFFFFu), w.n - 1)]);
        group_terms<DIM, G>(acc, c, dxi, dyi, dzi);
    }
}

// home_global's stencil (Physics.java:404-424) over the LDS cell table cst
// (cell c holds sorted positions cst[c] .. cst[c + 1] - 1).  Each visited
// run of cells goes through one call site of small_run (G = 1: one lane per
// home) or small_run_split (G lanes per home; lane sub of the group).
template <int DIM, bool DEBUG, bool MLDS, int G>
__device__ __forceinline__ void small_home(const ForceArgs &a, const WorldDev &w, const double *M,
                                           const uint32_t *SX, const uint32_t *SY, const uint32_t *SZ,
                                           const uint8_t *ST, const uint32_t *P, const uint16_t *cst,
                                           int hl, int sub, uint32_t hx, uint32_t hy, uint32_t hz,
                                           int row, Acc &acc)
{
    const int nx = w.nx;
    const bool wrap = w.wrap != 0;
    // stencil origin floor(x / cs) (Physics.java:404-405), not clamped
    const double ics = 1.0 / w.cs;
    const int cx0 = trunc_div_fast(hx, w.cs, ics);
    const int cy0 = trunc_div_fast(hy, w.cs, ics);
    const int cz0 = DIM == 3 ? trunc_div_fast(hz, w.cs, ics) : 0;
    // the row's three cells are consecutive keys, one run of the sorted order,
    // unless the row wraps in x: then cell by cell, with home_global's index
    // wrapping (twice in a periodic world: grids of one or two cells per axis
    // reach two cells past the edge) and range check
    const bool runs = cx0 >= 1 && cx0 <= nx - 2;
    const int nr = DIM == 3 ? 9 : 3;
    const int per = runs ? 1 : 3;
#pragma unroll 1
    for (int t = 0; t < nr * per; ++t) {
        const int r = t / per, ox = t % per - 1;
        int cy = wrap_cell(cy0 + r % 3 - 1, nx);
        int cz = DIM == 3 ? wrap_cell(cz0 + r / 3 - 1, nx) : 0;
        int lo, hi;
        if (runs) {
            const int ci = cx0 - 1 + cy * nx + cz * nx * nx;
            lo = cst[ci];
            hi = cst[ci + 3];
        } else {
            int cx = wrap_cell(cx0 + ox, nx);
            if (wrap) {
                cx = wrap_cell(cx, nx);
                cy = wrap_cell(cy, nx);
                if (DIM == 3) cz = wrap_cell(cz, nx);
            } else if (cx < 0 || cx >= nx || cy < 0 || cy >= nx || (DIM == 3 && (cz < 0 || cz >= nx))) {
                continue;
            }
            const int ci = cx + cy * nx + cz * nx * nx;
            lo = cst[ci];
            hi = cst[ci + 1];
        }
        if constexpr (G == 1)
            small_run<DIM, DEBUG, MLDS>(a, w, wrap, M, SX, SY, SZ, ST, P, hl, hx, hy, hz, row, acc, lo, hi);
        else
            small_run_split<DIM, DEBUG, MLDS, G>(a, w, wrap, M, SX, SY, SZ, ST, P, sub, hx, hy, hz, row, acc,
                                                 lo, hi);
    }
}

// Stable LSD radix sort in LDS of count entries (key << 16 | index) on
// their kbits key bits, src and dst ping-ponging; returns the buffer that
// holds the result.  Each wave owns a contiguous range of the entries, and
// within a wave the entries of one digit are ranked by lane (ballots on the
// digit's bits), so equal keys keep their order.  cnt: kSmallWaves x 2^7 words.
__device__ __attribute__((noinline)) uint32_t *small_radix_sort(uint32_t *src, uint32_t *dst, int count,
                                                                int kbits, uint32_t *cnt, uint32_t *s_w)
{
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const int passes = (kbits + kSmallDigitBits - 1) / kSmallDigitBits;
    const int chunk = ((count + kSmallWaves - 1) / kSmallWaves + kWave - 1) & ~(kWave - 1);
    const int c0 = min(wv * chunk, count), c1 = min(c0 + chunk, count);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t *mycnt = cnt + wv * (1 << kSmallDigitBits);
    for (int p = 0; p < passes; ++p) {
        const int shift = 16 + p * kSmallDigitBits;
        const int db = min(kSmallDigitBits, kbits - p * kSmallDigitBits);
        const uint32_t dmask = (1u << db) - 1u;
        for (int t = tid; t < kSmallWaves << kSmallDigitBits; t += kSmallThreads) cnt[t] = 0u;
        __syncthreads();
        for (int i = c0 + lane; i < c1; i += kWave) atomicAdd(&mycnt[(src[i] >> shift) & dmask], 1u);
        __syncthreads();
        {
            // digit d of wave ww starts after every smaller digit and after
            // digit d of the waves before ww: one exclusive scan over the
            // counts in (digit, wave) order, two entries per thread
            const int e0 = 2 * tid, d0 = e0 / kSmallWaves, w0 = e0 % kSmallWaves;
            const uint32_t v0 = d0 < (1 << db) ? cnt[(w0 << kSmallDigitBits) + d0] : 0u;
            const uint32_t v1 = d0 < (1 << db) ? cnt[((w0 + 1) << kSmallDigitBits) + d0] : 0u;
            uint32_t total;
            const uint32_t base = block_excl_scan<kSmallThreads>(v0 + v1, s_w, total);
            if (d0 < (1 << db)) {
                cnt[(w0 << kSmallDigitBits) + d0] = base;
                cnt[((w0 + 1) << kSmallDigitBits) + d0] = base + v0;
            }
        }
        __syncthreads();
        for (int i0 = c0; i0 < c1; i0 += kWave) {
            const int i = i0 + lane;
            const bool valid = i < c1;
            const uint32_t e = valid ? src[i] : 0u;
            const uint32_t d = (e >> shift) & dmask;
            uint64_t peers = __ballot(valid);
            for (int b = 0; b < db; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t bb = __ballot(bit);
                peers &= bit ? bb : ~bb;
            }
            const uint32_t r = (uint32_t)__popcll(peers & lt);
            uint32_t base = 0;
            if (valid) base = mycnt[d];
            if (valid) dst[base + r] = e;
            // the wave's next entries of digit d follow this group
            if (valid && r == 0) mycnt[d] = base + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        uint32_t *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Cell table of count keys in non-decreasing order, key(j) = get(j) < nc:
// tab[c] = the first position whose key is >= c, for c in [0, nc].  Keys are
// bounded by nc again only to keep every LDS index inside the table.
template <typename Get>
__device__ __forceinline__ void small_table(uint16_t *tab, int count, int nc, Get get)
{
    if (count <= 0) {
        for (int c = threadIdx.x; c <= nc; c += kSmallThreads) tab[c] = 0;
        return;
    }
#pragma unroll 1
    for (int r0 = 0; r0 * kSmallThreads <= count; r0 += kSmallU) {
        int kj[kSmallU], kp[kSmallU];
#pragma unroll
        for (int u = 0; u < kSmallU; ++u) {
            const int j = threadIdx.x + (r0 + u) * kSmallThreads;
            kj[u] = get(min(j, count - 1));
            kp[u] = get(max(min(j, count) - 1, 0));
        }
#pragma unroll
        for (int u = 0; u < kSmallU; ++u) {
            const int j = threadIdx.x + (r0 + u) * kSmallThreads;
            if (j <= count) {
                const int hi = j < count ? min(kj[u], nc) : nc;
                const int lo = j > 0 ? min(kp[u], nc) : -1;
                for (int c = lo + 1; c <= hi; ++c) tab[c] = (uint16_t)j;
            }
        }
    }
}

// Stable sort of count <= kSmallThreads entries by key: each thread ranks
// its entry against all others (lower key, or equal key and lower position).
__device__ __forceinline__ void small_rank_sort(const uint32_t *src, uint32_t *dst, int count)
{
    const int t = threadIdx.x;
    if (t >= count) return;
    const uint32_t e = src[t], k = e >> 16;
    int r = 0;
    for (int q = 0; q < count; ++q) {
        const uint32_t kq = src[q] >> 16;
        r += (kq < k) | ((kq == k) & (q < t));
    }
    dst[r] = e;
}

template <int DIM, bool DEBUG, bool MLDS>
__global__ __launch_bounds__(kSmallThreads) void k_step_small(SmallArgs s)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t s_w[kSmallWaves];
    __shared__ unsigned long long s_q[kSmallWaves];
    __shared__ int s_last;
    const ForceArgs &a = s.f;
    const SmallLds L = small_lds_layout(DIM, MLDS ? a.ntypes_max : 0, s.n_max, s.cells_max);
    double *M = reinterpret_cast<double *>(smem + L.m);
    uint32_t *P0 = reinterpret_cast<uint32_t *>(smem + L.p0);
    uint32_t *P1 = reinterpret_cast<uint32_t *>(smem + L.p1);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(smem + L.cnt);
    uint16_t *cst = reinterpret_cast<uint16_t *>(smem + L.cst);
    uint32_t *SY = reinterpret_cast<uint32_t *>(smem + L.sy);
    uint32_t *SZ = reinterpret_cast<uint32_t *>(smem + L.sz);
    uint8_t *ST = smem + L.st;

    const int tid = threadIdx.x;
    const uint2 bw = a.blocks[blockIdx.x];
    const WorldDev w = a.worlds[bw.x];
    const int n = w.n, off = w.off, nx = w.nx;
    const int nc = DIM == 3 ? nx * nx * nx : nx * nx;
    const int first = (int)bw.y - off;           // the workgroup's first home, world-local
    const bool lead = first == 0;                // writes the world's sort by-products
    int kbits = 0;
    while ((1 << kbits) < nc) ++kbits;
    // rounds of kSmallL entries i = tid + (r0 + u) kSmallThreads cover [0, n)
    const int nrounds = (n + kSmallL * kSmallThreads - 1) / (kSmallL * kSmallThreads);

    if constexpr (MLDS)
        for (int t = tid; t < w.ntypes * w.ntypes; t += kSmallThreads) M[t] = a.matrix[w.moff + t] * w.R;

    // 1. cell keys in the pre-sort order (Physics.java:362-375), packed with
    //    the index, and the previous step's keys of the same slots (O, in P1)
    //    for the incremental sort
    uint16_t *O = reinterpret_cast<uint16_t *>(P1);
    bool old_ok = s.skeys_in != nullptr && n > 0;
    {
        bool bad = false;
        const double ics = 1.0 / w.cs;
#pragma unroll 1
        for (int g = 0; g < nrounds; ++g) {
            uint32_t px[kSmallL], py[kSmallL], pz[kSmallL], po[kSmallL];
#pragma unroll
            for (int u = 0; u < kSmallL; ++u) {
                const int i = off + min(tid + (g * kSmallL + u) * kSmallThreads, n - 1);
                px[u] = a.sx[i];
                py[u] = a.sy[i];
                pz[u] = DIM == 3 ? a.sz[i] : 0u;
                po[u] = old_ok ? s.skeys_in[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kSmallL; ++u) {
                const int i = tid + (g * kSmallL + u) * kSmallThreads;
                if (i < n) {
                    uint32_t key = (uint32_t)key_axis_fast(px[u], w.cs, ics, nx) +
                                   (uint32_t)key_axis_fast(py[u], w.cs, ics, nx) * (uint32_t)nx;
                    if constexpr (DIM == 3) key += (uint32_t)key_axis_fast(pz[u], w.cs, ics, nx) * (uint32_t)(nx * nx);
                    P0[i] = key << 16 | (uint32_t)i;
                    const uint32_t o = po[u] - w.cell_base;     // wraps to >= nc below the world's cells
                    bad |= o >= (uint32_t)nc;
                    O[i] = (uint16_t)min(o, 0xFFFFu);
                    if (lead) s.keys_out[off + i] = w.cell_base + key;
                }
            }
        }
        old_ok = !__syncthreads_or(bad) && old_ok;
    }

    // 2. the stable sort of (key, index), the reference's counting sort.
    //    Incremental path: the slots are the previous step's sorted order, so
    //    where the previous keys O are non-decreasing, the particles still in
    //    their cell ("stayers": key == O) already form a sorted subsequence and
    //    only the others ("movers") need sorting.  A particle of key c then
    //    lands at cst[c] + (stayers of key c before it) + (movers of key c
    //    before it); the stayers of key c are those of O's run of c.
    uint32_t *src = P0, *dst = P1;
    bool done = false;
    if (old_ok) {
        bool bad = false;
#pragma unroll 1
        for (int g = 0; g < nrounds; ++g) {
#pragma unroll
            for (int u = 0; u < kSmallL; ++u) {
                const int i = tid + (g * kSmallL + u) * kSmallThreads;
                const int ic = max(min(i, n - 1), 1);
                bad |= (i >= 1) & (i < n) & (O[ic] < O[ic - 1]);
            }
        }
        old_ok = !__syncthreads_or(bad);
    }
    if (old_ok) {
        // stayer prefix Sp[i] (i in [0, n]) over contiguous chunks of ch <=
        // kSmallPer entries per thread; movers compacted in index order at
        // i - Sp[i]
        uint16_t *Sp = reinterpret_cast<uint16_t *>(SY);
        uint32_t *MV = reinterpret_cast<uint32_t *>(smem + ((L.sy + (size_t)(n + 1) * 2 + 3) & ~(size_t)3));
        const int mmax = n / 4;
        uint32_t *MV2 = MV + mmax;
        const int ch = (n + kSmallThreads - 1) / kSmallThreads;
        const int i0 = min(tid * ch, n), i1 = min(i0 + ch, n);
        uint32_t stay = 0;
#pragma unroll 1
        for (int r0 = 0; r0 < ch; r0 += kSmallU) {
            uint32_t e[kSmallU];
            uint16_t o[kSmallU];
#pragma unroll
            for (int u = 0; u < kSmallU; ++u) {
                const int ic = min(i0 + r0 + u, n - 1);
                e[u] = P0[ic];
                o[u] = O[ic];
            }
#pragma unroll
            for (int u = 0; u < kSmallU; ++u) stay += (i0 + r0 + u < i1) & ((e[u] >> 16) == o[u]);
        }
        uint32_t total;
        uint32_t sp = block_excl_scan<kSmallThreads>(stay, s_w, total);
        const int m = n - (int)total;
        if (m <= mmax) {
#pragma unroll 1
            for (int i = i0; i < i1; ++i) {
                const uint32_t e = P0[i];
                Sp[i] = (uint16_t)sp;
                if ((e >> 16) == O[i]) ++sp;
                else MV[i - (int)sp] = e;
            }
            if (tid == 0) Sp[n] = (uint16_t)total;
            __syncthreads();
            uint32_t *mv = MV2;
            if (m <= kSmallThreads) small_rank_sort(MV, MV2, m);
            else mv = small_radix_sort(MV, MV2, m, kbits, cnt, s_w);
            __syncthreads();
            // A mover of sorted rank q, key c and slot i lands after S_q
            // stayers -- those before slot i clamped to O's run of c: every
            // stayer of a smaller key, or of key c and a smaller slot -- and
            // after the q movers before it.  Positions first: O shares P1 with
            // the output.  The two searches of O's run go side by side.
            uint32_t *Sq = mv == MV2 ? MV : MV2;     // S_q, non-decreasing in q
            int mpos[kSmallPer];
#pragma unroll
            for (int r = 0; r < kSmallPer; ++r) {
                const int q = tid + r * kSmallThreads;
                mpos[r] = 0;
                if (q < m) {
                    const uint32_t ee = mv[q];
                    const int c = (int)(ee >> 16), i = (int)(ee & 0xFFFFu);
                    int lo = 0, hi = n, lo2 = 0, hi2 = n;    // first slot with O >= c; with O > c
                    while (lo < hi || lo2 < hi2) {
                        const int mid = (lo + hi) >> 1, mid2 = (lo2 + hi2) >> 1;
                        const int o1 = (int)O[min(mid, n - 1)], o2 = (int)O[min(mid2, n - 1)];
                        if (lo < hi) {
                            if (o1 < c) lo = mid + 1;
                            else hi = mid;
                        }
                        if (lo2 < hi2) {
                            if (o2 <= c) lo2 = mid2 + 1;
                            else hi2 = mid2;
                        }
                    }
                    const int sq = (int)Sp[min(max(i, lo), lo2)];
                    Sq[q] = (uint32_t)sq;
                    mpos[r] = sq + q;
                }
            }
            __syncthreads();
            // A stayer of stayer rank k (= Sp[i] at its slot i) lands after the
            // k stayers before it and the movers with S_q <= k.  Each thread
            // places the stayers of its chunk [i0, i1) in slot order, counting
            // those movers as k grows from a search at the chunk's first rank.
            {
                const int k0 = (int)Sp[i0];
                int cm = 0, chi = m;
                while (cm < chi) {
                    const int mid = (cm + chi) >> 1;
                    if ((int)Sq[mid] <= k0) cm = mid + 1;
                    else chi = mid;
                }
#pragma unroll 1
                for (int r0 = 0; r0 < ch; r0 += kSmallU) {
                    uint32_t e[kSmallU];
                    int sp[kSmallU + 1];
#pragma unroll
                    for (int u = 0; u <= kSmallU; ++u) sp[u] = (int)Sp[min(i0 + r0 + u, n)];
#pragma unroll
                    for (int u = 0; u < kSmallU; ++u) e[u] = P0[min(i0 + r0 + u, n - 1)];
#pragma unroll
                    for (int u = 0; u < kSmallU; ++u) {
                        if (i0 + r0 + u < i1 && sp[u + 1] != sp[u]) {
                            while (cm < m && (int)Sq[cm] <= sp[u]) ++cm;
                            P1[sp[u] + cm] = e[u];
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kSmallPer; ++r) {
                const int q = tid + r * kSmallThreads;
                if (q < m) P1[mpos[r]] = mv[q];
            }
            __syncthreads();
            src = P1;
            dst = P0;
            done = true;
        } else {
            __syncthreads();
        }
    }
    if (!done) {
        src = small_radix_sort(P0, P1, n, kbits, cnt, s_w);
        dst = src == P0 ? P1 : P0;
    }
    uint32_t *SX = dst;                          // the free buffer

    // 3. the cell table of the sorted order
    small_table(cst, n, nc, [&](int j) { return (int)(src[j] >> 16); });
    __syncthreads();                             // the incremental sort's scratch is read no more
    // 4. sorted positions and types (the gather of Physics.java:122-131's
    //    particle copy) and the sort's by-products
#pragma unroll 1
    for (int g = 0; g < nrounds; ++g) {
        uint32_t gx[kSmallL], gy[kSmallL], gz[kSmallL], e[kSmallL];
        int gt[kSmallL];
#pragma unroll
        for (int u = 0; u < kSmallL; ++u) e[u] = src[min(tid + (g * kSmallL + u) * kSmallThreads, n - 1)];
#pragma unroll
        for (int u = 0; u < kSmallL; ++u) {
            const int go = off + min((int)(e[u] & 0xFFFFu), n - 1);
            gx[u] = a.sx[go];
            gy[u] = a.sy[go];
            gz[u] = DIM == 3 ? a.sz[go] : 0u;
            gt[u] = a.type[go];
        }
#pragma unroll
        for (int u = 0; u < kSmallL; ++u) {
            const int j = tid + (g * kSmallL + u) * kSmallThreads;
            if (j < n) {
                SX[j] = gx[u];
                SY[j] = gy[u];
                if constexpr (DIM == 3) SZ[j] = gz[u];
                ST[j] = (uint8_t)gt[u];
                if (lead) {
                    s.perm_out[off + j] = (uint32_t)off + min(e[u] & 0xFFFFu, (uint32_t)(n - 1));
                    s.skeys_out[off + j] = w.cell_base + (e[u] >> 16);
                }
            }
        }
    }
    __syncthreads();
    unsigned long long q = 0;
    if (lead) {
        for (int c = tid; c < nc; c += kSmallThreads) {
            const uint32_t b = cst[c], e = cst[c + 1];
            s.cell_start_out[w.cell_base + c] = (uint32_t)off + b;
            s.cell_end_out[w.cell_base + c] = (uint32_t)off + e;
            q += (unsigned long long)(e - b) * (e - b);
        }
        q = block_sum_u64<kSmallThreads>(q, s_q);
    }

    // 5. force and integration of the workgroup's homes [first, first + homes):
    //    G lanes per home where the homes leave threads free (G = 16 for at
    //    most 64 homes, 4 for at most 256), else kSmallThreads homes at a time
    //    (a workgroup may own a whole world of a batch, so that the world is
    //    sorted once)
    const int hend = min(first + s.homes, n);
    unsigned long long mine = 0;
    const int gsh = s.homes <= kSmallThreads / 16 ? 4 : s.homes <= kSmallThreads / 4 ? 2 : 0;   // log2 G
    if (gsh > 0) {
        const int hl = first + (tid >> gsh), sub = tid & ((1 << gsh) - 1);
        if (hl < hend) {                         // the whole group of the home
            Acc acc;
            acc.x = acc.y = acc.z = 0.0;
            acc.cnt = 0;
            acc.hash = 0ull;
            const uint32_t hx = SX[hl], hy = SY[hl], hz = DIM == 3 ? SZ[hl] : 0u;
            const int htype = ST[hl];
            if (gsh == 4)
                small_home<DIM, DEBUG, MLDS, 16>(a, w, M, SX, SY, SZ, ST, src, cst, hl, sub, hx, hy, hz,
                                                 htype * w.ntypes, acc);
            else
                small_home<DIM, DEBUG, MLDS, 4>(a, w, M, SX, SY, SZ, ST, src, cst, hl, sub, hx, hy, hz,
                                                htype * w.ntypes, acc);
            for (int d = 1; d < (1 << gsh); d <<= 1) {
                acc.cnt += __shfl_xor(acc.cnt, d, kWave);
                if constexpr (DEBUG) acc.hash += __shfl_xor(acc.hash, d, kWave);
            }
            if (sub == 0) {
                const uint32_t ho = (uint32_t)(off + min((int)(src[hl] & 0xFFFFu), n - 1));
                if (w.wrap) finish_home<DIM, true, DEBUG>(a, w, acc, hx, hy, hz, htype, ho, off + hl);
                else finish_home<DIM, false, DEBUG>(a, w, acc, hx, hy, hz, htype, ho, off + hl);
                mine += (unsigned long long)acc.cnt;
            }
        }
    } else {
#pragma unroll 1
        for (int h0 = first; h0 < hend; h0 += kSmallThreads) {
            const int hl = h0 + tid;
            if (hl < hend) {
                Acc acc;
                acc.x = acc.y = acc.z = 0.0;
                acc.cnt = 0;
                acc.hash = 0ull;
                const uint32_t hx = SX[hl], hy = SY[hl], hz = DIM == 3 ? SZ[hl] : 0u;
                const int htype = ST[hl];
                const uint32_t ho = (uint32_t)(off + min((int)(src[hl] & 0xFFFFu), n - 1));
                small_home<DIM, DEBUG, MLDS, 1>(a, w, M, SX, SY, SZ, ST, src, cst, hl, 0, hx, hy, hz,
                                                htype * w.ntypes, acc);
                if (w.wrap) finish_home<DIM, true, DEBUG>(a, w, acc, hx, hy, hz, htype, ho, off + hl);
                else finish_home<DIM, false, DEBUG>(a, w, acc, hx, hy, hz, htype, ho, off + hl);
                mine += (unsigned long long)acc.cnt;
            }
        }
    }
    const unsigned long long tot = block_sum_u64<kSmallThreads>(mine, s_q);
    // 6. counters.  Agent-scope atomics on one location are coherent across
    //    the XCDs; the ticket's increment depends on the values returned by
    //    this workgroup's adds, so it is issued only after they are performed,
    //    and the last workgroup's exchanges read every workgroup's adds.
    if (tid == 0) {
        unsigned long long dep = 0;
        if (q) dep += atomicAdd(s.acc + 1, q);
        if (tot) {
            dep += atomicAdd(s.acc, tot);
            dep += atomicAdd(s.acc + 2 + bw.x, tot);
        }
        const unsigned long long ticket = atomicAdd(s.acc + s.ncounters, dep == ~0ull ? 2ull : 1ull);
        s_last = ticket == (unsigned long long)(gridDim.x - 1);
    }
    __syncthreads();
    // the last workgroup publishes the step's counters and clears the accumulator
    if (s_last) {
        for (int t = tid; t <= s.ncounters; t += kSmallThreads) {
            const unsigned long long v = t < s.ncounters ? atomicExch(s.acc + t, 0ull) : s.tag;
            s.counters[t] = v;
            if (s.stat_out) s.stat_out[t] = v;
        }
        __syncthreads();
        if (tid == 0) atomicExch(s.acc + s.ncounters, 0ull);
    }
}

}  // namespace plife
