// plife_force.hpp -- the fused velocity + position kernel of the particle-life step.
//
// Reference: backend/Physics.java:397-450 (updateVelocity, updatePosition),
// the accelerator lambda app/Main.java:275-280, Range.java:40-96.
//
// Work decomposition: a workgroup of BS threads owns BS consecutive home
// particles of the cell-sorted order, so every wave is full.  It stages, for
// each grid row it touches (2D: row = cy, 3D: row = cy + cz*nx), the
// 3^(DIM-1) neighbour rows x (span + 2) cells of the sorted SoA in LDS; each
// thread owns one home and, per stencil row, scans the three staged cells
// around its own column -- exactly the cells the reference's 3x3 (3x3x3)
// stencil visits.
//
// Each home particle runs two phases over its candidates:
//   A  FP32 screen: exact int32 minimum-image deltas (positions are 32-bit
//      fixed point), |D|^2 in FP32, keep 0 < |D|^2 <= rmax^2 (1 + 2^-20).  Kept
//      candidate indices are appended to a per-thread list in LDS.
//   B  For listed candidates only: the reference's predicate in FP64
//      (d2 = dx*dx + dy*dy, d2 != 0 && d2 <= rmax*rmax; Physics.java:430-432) and
//      the force law in FP64 (FP64 FMA issues at the non-packed FP32 rate on
//      CDNA4), accumulated in FP64.
// The screen is a strict superset of the FP64 predicate (its FP32 rounding is
// below 2^-21 relative), so neighbour sets are exactly the reference's.
#pragma once

#include <type_traits>

#include "plife_kernels.hpp"

namespace plife {

struct ForceArgs {
    // step-start state in the pre-sort order, read through the permutation
    const uint32_t *perm;            // sorted slot -> pre-sort index
    const uint32_t *sx, *sy, *sz;
    const float *vx, *vy, *vz;
    const int32_t *type, *id;
    // the new state, written in the sorted order (Physics.particles after the swap)
    uint32_t *ox, *oy, *oz;
    float *ovx, *ovy, *ovz;
    int32_t *otype, *oid;
    const uint32_t *skeys;           // batch cell key of each sorted slot
    const uint32_t *cell_start, *cell_end;
    const double *matrix;            // per world ntypes*ntypes at moff, row = acted-on type
    const WorldDev *worlds;
    const uint2 *blocks;             // per workgroup: (world, first home slot)
    int block_base;                  // this launch's first entry of blocks
    int ntypes_max, m_in_lds;
    int cap, lcap, max_cells;
    int wpool;                       // mask words per workgroup of the balanced path (0: windowed only)
    unsigned long long *pair_total;  // interacting pairs of the step, all worlds
    unsigned long long *pair_world;  // per world
    uint32_t *block_work;            // pinned host memory: per table entry its pairs (saturated); [nb_total] = tag
    uint32_t work_tag;               // identifies the block table the counts belong to
    int nb_total;
    int32_t *dbg_count;
    unsigned long long *dbg_hash;
    float *dbg_acc;                  // debug: the step's acceleration per sorted slot, (x, y, z)
};

struct Acc {
    double x, y, z;
    int cnt;
    unsigned long long hash;
};

__device__ __forceinline__ unsigned long long id_hash(uint32_t id)
{
    unsigned long long z = (unsigned long long)id + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One pair in FP64, in units of 2^-32 (the fixed-point grid), where scaling
// by powers of two is exact: with D = x_q - x_p, rel = D/rmax, dist = |rel|,
// the reference adds rel*(f/dist)*(rmax*force*dt) = D*(f/dist)*(force*dt) where
//   f/dist = 1/beta - 1/dist                             (dist <  beta)
//          = a*(1 - |1 + beta - 2 dist|/(1 - beta))/dist (dist >= beta)
// (app/Main.java:275-280).  acc accumulates D*(f/dist) in units of 2^-32.
// With u = |D| = dist R (R = rmax 2^32) and y = 1/|D|: 1/dist = R y, and the
// attractive branch is (a R) (1 - |1 + beta - 2 S u|/(1 - beta)) y, so the
// caller passes mR = a * R (staged that way with the matrix).
template <int DIM>
__device__ __forceinline__ double pair_d2(double dx, double dy, double dz)
{
#pragma clang fp contract(off)
    // JOML lengthSquared without FMA: x*x + (y*y + z*z) (exact 2^64 scaling)
    return DIM == 3 ? dx * dx + (dy * dy + dz * dz) : dx * dx + dy * dy;
}

// f/dist of the force law for a pair in range (0 < d2 <= rmax^2).
__device__ __forceinline__ double pair_coef(double d2, double mR, const WorldDev &a)
{
    const double y0 = (double)__builtin_amdgcn_rsqf((float)d2);
    const double e = d2 * y0;
    const double h = fma(-e, y0, 1.0);
    const double y = fma(0.5 * y0, h, y0);        // 1/|D|, one Newton step from FP32
    const double u = d2 * y;                       // |D| = dist R
    const double t = fabs(fma(u, a.m2S, a.one_plus_beta));
    const double g = fma(t, -a.inv_1mb, 1.0);
    const double att = (mR * g) * y;
    const double rep = fma(-a.R, y, a.inv_beta);   // 1/beta - 1/dist
    return u < a.betaR ? rep : att;
}

template <int DIM, bool DEBUG>
__device__ __forceinline__ void pair64(Acc &acc, int dxi, int dyi, int dzi, double mR, uint32_t qid,
                                       const WorldDev &a)
{
    const double dx = (double)dxi, dy = (double)dyi, dz = (double)dzi;
    const double d2 = pair_d2<DIM>(dx, dy, dz);
    const bool in = (d2 != 0.0) & (d2 <= a.r2u);     // Physics.java:432
    if (in) {
        const double c = pair_coef(d2, mR, a);
        acc.x = fma(c, dx, acc.x);
        acc.y = fma(c, dy, acc.y);
        if constexpr (DIM == 3) acc.z = fma(c, dz, acc.z);
        acc.cnt += 1;
        if constexpr (DEBUG) acc.hash += id_hash(qid);
    }
}

// pair64 with no branch: the coefficient of a pair out of range (or
// rejected by the caller, ok = false) is selected to 0, which leaves the sums
// bit-identical (a sum that starts at +0 is never -0), so that the chains of
// neighbouring candidates can overlap.
template <int DIM, bool DEBUG>
__device__ __forceinline__ void pair64_sel(Acc &acc, bool ok, int dxi, int dyi, int dzi, double mR,
                                           uint32_t qid, const WorldDev &a)
{
    const double dx = (double)dxi, dy = (double)dyi, dz = (double)dzi;
    const double d2 = pair_d2<DIM>(dx, dy, dz);
    const bool in = ok & (d2 != 0.0) & (d2 <= a.r2u);    // Physics.java:432
    double cf = pair_coef(d2, mR, a);
    asm volatile("" : "+v"(cf));        // computed for every candidate: no branch
    const double c = in ? cf : 0.0;
    acc.x = fma(c, dx, acc.x);
    acc.y = fma(c, dy, acc.y);
    if constexpr (DIM == 3) acc.z = fma(c, dz, acc.z);
    acc.cnt += in ? 1 : 0;
    if constexpr (DEBUG)
        if (in) acc.hash += id_hash(qid);
}

// The block's pairs, for the host to order the next updates' blocks
// heaviest first; entry nb_total tags the table they belong to.
__device__ __forceinline__ void record_block_work(const ForceArgs &a, int blk,
                                                  unsigned long long pairs)
{
    const int e = a.block_base + blk;
    a.block_work[e] = (uint32_t)min(pairs, 0xFFFFFFFFull);
    if (e == 0) a.block_work[a.nb_total] = a.work_tag;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, kWave));
    return v;
}

// Phase A screen on exact fixed-point deltas.  In a clamped world the int32
// difference may wrap across the box (|D| >= 1/2 > rmax); phase B's exact
// sign test drops those, so the screen stays a superset.
template <int DIM>
__device__ __forceinline__ bool screen(uint32_t hx, uint32_t hy, uint32_t hz, uint32_t qx,
                                       uint32_t qy, uint32_t qz, float r2hi)
{
    const float dx = (float)(int)(qx - hx), dy = (float)(int)(qy - hy);
    float d2 = dx * dx;
    d2 = fmaf(dy, dy, d2);
    if constexpr (DIM == 3) {
        const float dz = (float)(int)(qz - hz);
        d2 = fmaf(dz, dz, d2);
    }
    // d2 == 0 (the home itself, coincident particles) passes the screen; the
    // FP64 predicate of phase B drops it
    return d2 <= r2hi;
}

// Periodic distance in cells between columns (rows, layers) a and b of n.
__device__ __forceinline__ int periodic_cells(int a, int b, int n)
{
    int d = a - b;
    d = d < 0 ? -d : d;
    d %= n;
    return min(d, n - d);
}

// 16-bit relative coordinate of fixed-point p against the origin o (floor of
// the periodic int32 difference over 2^sh), as the low 16 bits of a u32.
__device__ __forceinline__ uint32_t rel16(uint32_t p, uint32_t o, int sh)
{
    return (uint32_t)((int)(p - o) >> sh) & 0xFFFFu;
}

typedef short short2v __attribute__((ext_vector_type(2)));

// Phase A screen on packed 16-bit relative coordinates (x | y << 16, z apart):
// packed subtract and a saturating dot product against the quantised bound.
// Each coordinate is at most one unit off its exact quotient, so the bound
// (Rq + 2)^2 keeps every pair with |D| <= rmax (1 + 2^-20).
template <int DIM>
__device__ __forceinline__ bool screen16(uint32_t q, uint32_t h, uint32_t qz, uint32_t hz, int bq)
{
    const short2v d = __builtin_bit_cast(short2v, q) - __builtin_bit_cast(short2v, h);
    int zz = 0;
    if constexpr (DIM == 3) {
        const int dz = (int)(short)(uint16_t)(qz - hz);
        zz = dz * dz;
    }
    return __builtin_amdgcn_sdot2(d, d, zz, true) <= bq;
}

__device__ __forceinline__ int wrap_cell(int c, int n)
{
    return c < 0 ? c + n : (c >= n ? c - n : c);
}

// FP64 stencil origin floor(x / cs) (Physics.java:404-405), not clamped.
__device__ __forceinline__ int stencil_origin(uint32_t k, double cs)
{
    return (int)floor(((double)k * kInvFix) / cs);
}

// Physics.java:443-450 + Range.wrap / Range.clamp in FP64, rounded to 2^-32.
__device__ __forceinline__ uint32_t integrate_axis(uint32_t k, double v, double dt, bool wrap)
{
#pragma clang fp contract(off)
    double x = v * dt + (double)k * kInvFix;    // JOML mulAdd: v*dt + x
    if (!(x == x)) return 0u;
    if (wrap) {
        x = x - floor(x);                       // equals Range.wrap's repeated +-1
        const double s = rint(x * kFix);        // [0, 2^32]; 2^32 (x == 1.0) wraps to 0
        return (uint32_t)(unsigned long long)s;
    }
    x = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
    const double s = rint(x * kFix);
    return s >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)(unsigned long long)s;
}

template <int DIM, bool WRAP, bool DEBUG>
__device__ __forceinline__ void finish_home(const ForceArgs &a, const WorldDev &w, const Acc &acc,
                                            uint32_t hx, uint32_t hy, uint32_t hz, int htype,
                                            uint32_t o, int g)
{
    const double nvx = w.ff * (double)a.vx[o] + w.kdt * acc.x;
    const double nvy = w.ff * (double)a.vy[o] + w.kdt * acc.y;
    a.ovx[g] = (float)nvx;
    a.ovy[g] = (float)nvy;
    a.ox[g] = integrate_axis(hx, nvx, w.dt, WRAP);
    a.oy[g] = integrate_axis(hy, nvy, w.dt, WRAP);
    if constexpr (DIM == 3) {
        const double nvz = w.ff * (double)a.vz[o] + w.kdt * acc.z;
        a.ovz[g] = (float)nvz;
        a.oz[g] = integrate_axis(hz, nvz, w.dt, WRAP);
    }
    a.otype[g] = htype;
    a.oid[g] = a.id[o];
    if constexpr (DEBUG) {
        a.dbg_count[g] = acc.cnt;
        a.dbg_hash[g] = acc.hash;
        // the force law's acceleration (Physics.java:437 summed, over dt) in FP32
        a.dbg_acc[3 * (size_t)g + 0] = (float)(w.ka * acc.x);
        a.dbg_acc[3 * (size_t)g + 1] = (float)(w.ka * acc.y);
        a.dbg_acc[3 * (size_t)g + 2] = DIM == 3 ? (float)(w.ka * acc.z) : 0.0f;
    }
}

// Candidates lo..hi-1 of the cell-sorted order for home g (slot), FP64
// throughout: the body of the reference's per-cell loop (Physics.java:413-424).
template <int DIM, bool WRAP, bool DEBUG, bool MLDS>
__device__ __forceinline__ void visit_run(const ForceArgs &a, const WorldDev &w, const double *Mlds,
                                          int g, uint32_t hx, uint32_t hy, uint32_t hz, int row,
                                          Acc &acc, int lo, int hi)
{
    for (int j = lo; j < hi; ++j) {
        if (j == g) continue;
        const uint32_t jo = a.perm[j];
        const uint32_t qx = a.sx[jo], qy = a.sy[jo], qz = DIM == 3 ? a.sz[jo] : 0u;
        const int dxi = (int)(qx - hx), dyi = (int)(qy - hy);
        const int dzi = DIM == 3 ? (int)(qz - hz) : 0;
        if (!WRAP) {
            // plain difference of the clamped world: skip |D| >= 1/2
            bool ok = ((dxi >= 0) == (qx >= hx)) & ((dyi >= 0) == (qy >= hy));
            if (DIM == 3) ok &= ((dzi >= 0) == (qz >= hz));
            if (!ok) continue;
        }
        const int idx = row + a.type[jo];
        const double m = MLDS ? Mlds[idx] : a.matrix[w.moff + idx] * w.R;
        pair64<DIM, DEBUG>(acc, dxi, dyi, dzi, m, DEBUG ? (uint32_t)a.id[jo] : 0u, w);
    }
}

// Literal reference stencil over 3^DIM cells straight from global memory
// (Physics.java:404-424), FP64 throughout.  Used for strips whose staged
// neighbourhood does not fit in LDS, for sparse worlds and for grids with
// fewer than 3 cells per axis (where the reference visits repeated cells).
// Where a stencil row does not wrap in x its three cells are consecutive keys,
// so the row is one run of the cell-sorted order: the same candidates in the
// same order as the per-cell loop, from two bound loads instead of six, all
// issued before the first candidate is visited.
template <int DIM, bool WRAP, bool DEBUG, bool MLDS>
__device__ void home_global(const ForceArgs &a, const WorldDev &w, const double *Mlds, int g,
                            uint32_t hx, uint32_t hy, uint32_t hz, int row, Acc &acc)
{
    constexpr int NR = DIM == 3 ? 9 : 3;
    const int nx = w.nx;
    const int cx0 = stencil_origin(hx, w.cs);
    const int cy0 = stencil_origin(hy, w.cs);
    const int cz0 = DIM == 3 ? stencil_origin(hz, w.cs) : 0;
    if (cx0 >= 1 && cx0 <= nx - 2) {
        int lo[NR], hi[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int cy = wrap_cell(cy0 + r % 3 - 1, nx);
            const int cz = DIM == 3 ? wrap_cell(cz0 + r / 3 - 1, nx) : 0;
            const uint32_t ci = w.cell_base + (uint32_t)(cx0 - 1 + cy * nx + cz * nx * nx);
            lo[r] = (int)a.cell_start[ci];
            hi[r] = (int)a.cell_end[ci + 2];
        }
#pragma unroll
        for (int r = 0; r < NR; ++r)
            visit_run<DIM, WRAP, DEBUG, MLDS>(a, w, Mlds, g, hx, hy, hz, row, acc, lo[r], hi[r]);
        return;
    }
    const int zlo = DIM == 3 ? -1 : 0, zhi = DIM == 3 ? 1 : 0;
    for (int oz = zlo; oz <= zhi; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                int cx = wrap_cell(cx0 + ox, nx), cy = wrap_cell(cy0 + oy, nx);
                int cz = DIM == 3 ? wrap_cell(cz0 + oz, nx) : 0;
                if (WRAP) {
                    cx = wrap_cell(cx, nx);
                    cy = wrap_cell(cy, nx);
                    if (DIM == 3) cz = wrap_cell(cz, nx);
                } else if (cx < 0 || cx >= nx || cy < 0 || cy >= nx ||
                           (DIM == 3 && (cz < 0 || cz >= nx))) {
                    continue;
                }
                const uint32_t ci = w.cell_base + (uint32_t)(cx + cy * nx + cz * nx * nx);
                visit_run<DIM, WRAP, DEBUG, MLDS>(a, w, Mlds, g, hx, hy, hz, row, acc,
                                                  (int)a.cell_start[ci], (int)a.cell_end[ci]);
            }
}

// A workgroup of BS threads (BS/64 waves) owns BS consecutive home particles
// of the cell-sorted order, so every wave is full.  The homes span a few
// consecutive cells of one grid row, or (at row ends) of a few rows: each
// distinct row is a "part" whose 3^(DIM-1) stencil rows x (span + 2) cells are
// staged in LDS for the whole workgroup, in windows of at most `cap` particles
// when the neighbourhood is denser than the staging buffer.  Workgroups with
// more parts or staged cells than the cell table holds take the global path.
constexpr int kMaxParts = 4;
constexpr int kScreenPad = 8;    // >= the screen's group size U
constexpr int kDealKeys = 5;     // 4 position groups + the inactive threads of a short block

// Position group of a home inside its cell: y quartile (2D) or y/z halves (3D).
template <int DIM>
__device__ __forceinline__ int deal_key(uint32_t y, uint32_t z, int hrow, int nx, double cs)
{
    const float inv = (float)(kInvFix / cs);
    const int cy = hrow % nx;
    const float fy = (float)y * inv - (float)cy;           // ~[0, 1) within the cell
    if constexpr (DIM == 2) return min(3, max(0, (int)(fy * 4.0f)));
    const float fz = (float)z * inv - (float)(hrow / nx);
    return min(1, max(0, (int)(fz * 2.0f))) * 2 + min(1, max(0, (int)(fy * 2.0f)));
}

struct ForceLds {
    size_t m, sxy, sq, lists, wbase, sz, sqz, sid, off, gs, st, total;
};

// The candidate lists (windowed path: lcap u16 entries per thread) and the
// mask words (balanced path: a pool of wpool u32 masks + wpool u16 first
// indices, each home its own run of them) share one region; a workgroup takes
// one path.
__host__ __device__ inline ForceLds force_lds_layout(int dim, bool debug, int ntypes, bool m_in_lds,
                                                     int cap, int lcap, int bs, int max_cells,
                                                     int wpool)
{
    ForceLds L;
    size_t o = 0;
    L.m = o;     o += m_in_lds ? (size_t)ntypes * ntypes * 8 : 0;
    // the screened arrays carry kScreenPad spare entries: a screen step reads
    // a group of U candidates from one base, up to U - 1 past a lane's range
    L.sxy = o;   o += (size_t)(cap + kScreenPad) * 8;
    L.sq = o;    o += dim == 3 ? (size_t)(cap + kScreenPad) * 4 : 0;     // 16-bit screen (3D only)
    L.lists = o;
    L.wbase = o + (size_t)wpool * 4;
    {
        const size_t lsz = (size_t)lcap * bs * 2, wsz = (size_t)wpool * 6;
        o += lsz > wsz ? lsz : wsz;
    }
    o = (o + 3) & ~(size_t)3;
    L.sz = o;    o += dim == 3 ? (size_t)(cap + kScreenPad) * 4 : 0;
    L.sqz = o;   o += dim == 3 ? (size_t)(cap + kScreenPad) * 2 : 0;
    o = (o + 3) & ~(size_t)3;
    L.sid = o;   o += debug ? (size_t)cap * 4 : 0;
    L.off = o;   o += (size_t)(max_cells + 1) * 4;
    L.gs = o;    o += (size_t)max_cells * 4;
    L.st = o;    o += (size_t)cap;
    L.total = (o + 15) & ~(size_t)15;
    return L;
}

// MLDS: the type matrix is staged in LDS (ntypes <= 32); otherwise it is read
// from global memory.  A compile-time switch, so the hot loop's matrix load is
// a plain ds_read and never a generic (flat) load.
template <int DIM, bool WRAP, bool DEBUG, bool MLDS, int BS>
__global__ __launch_bounds__(BS) void k_force(ForceArgs a)
{
    constexpr int NR = DIM == 3 ? 9 : 3;      // stencil rows per part
    constexpr int RMID = DIM == 3 ? 4 : 1;    // the stencil row of the home's own row (oy = oz = 0)
    constexpr int U = 8;                      // screen group; lists keep U slots free (lcap >= U)
    constexpr int NW = BS / kWave;
    constexpr int PER = 256 / BS;             // count bins per thread in the balanced deal
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_part[kMaxParts][3];      // row, first column, last column
    __shared__ uint32_t s_cnt[NW + 1];
    __shared__ unsigned long long s_pairs[NW];
    __shared__ uint32_t s_qc[NW][kDealKeys];
    __shared__ uint32_t s_deal[BS];
    // balanced path: per home its staged index, kept candidates and mask words
    __shared__ uint16_t s_si[BS], s_hc[BS], s_wo[BS];
    __shared__ uint32_t s_bin[256];

    const ForceLds L = force_lds_layout(DIM, DEBUG, a.ntypes_max, MLDS, a.cap, a.lcap, BS, a.max_cells,
                                        a.wpool);
    double *M = reinterpret_cast<double *>(smem + L.m);
    uint2 *sxy = reinterpret_cast<uint2 *>(smem + L.sxy);
    uint32_t *sq = reinterpret_cast<uint32_t *>(smem + L.sq);
    uint16_t *sqz = reinterpret_cast<uint16_t *>(smem + L.sqz);
    uint16_t *lists = reinterpret_cast<uint16_t *>(smem + L.lists);
    uint32_t *wmask = reinterpret_cast<uint32_t *>(smem + L.lists);     // balanced path: same region
    uint16_t *wbase = reinterpret_cast<uint16_t *>(smem + L.wbase);
    uint32_t *szz = reinterpret_cast<uint32_t *>(smem + L.sz);
    uint32_t *sid = reinterpret_cast<uint32_t *>(smem + L.sid);
    uint32_t *off = reinterpret_cast<uint32_t *>(smem + L.off);
    uint32_t *gs = reinterpret_cast<uint32_t *>(smem + L.gs);
    uint8_t *st = reinterpret_cast<uint8_t *>(smem + L.st);

    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint2 bw = a.blocks[a.block_base + blk];
    const WorldDev w = a.worlds[bw.x];
    const int nx = w.nx;
    const uint32_t cb = w.cell_base;
    const int g0 = (int)bw.y;
    const int nact = min(BS, w.off + w.n - g0);

    if constexpr (MLDS)
        for (int t = tid; t < w.ntypes * w.ntypes; t += BS) M[t] = a.matrix[w.moff + t] * w.R;

    // Thread tid first stands for home g0 + tid (sorted order), to find the
    // parts: grid rows (2D: cy; 3D: cy + cz*nx) and columns of the homes.
    // Its position and type, and the origin's (home g0) position, are read
    // up front: independent loads whose latency overlaps the setup.
    const bool active0 = tid < nact;
    const int gs0 = g0 + (active0 ? tid : nact - 1);
    const uint32_t ho0 = a.perm[gs0];
    const uint32_t hx0 = a.sx[ho0], hy0 = a.sy[ho0], hz0 = DIM == 3 ? a.sz[ho0] : 0u;
    const int htype0 = a.type[ho0];
    const uint32_t go = a.perm[g0];
    const uint32_t ox = a.sx[go], oy = a.sy[go], oz = DIM == 3 ? a.sz[go] : 0u;
    const uint32_t hkey0 = a.skeys[gs0] - cb;
    const int hrow0 = (int)(hkey0 / (uint32_t)nx);
    const int hcx0 = (int)(hkey0 - (uint32_t)hrow0 * (uint32_t)nx);

    // parts = distinct rows among the homes (rows are non-decreasing with tid);
    // only active threads past the first read the previous home's key, so the
    // read stays inside [g0, g0 + nact)
    const int prow = (tid > 0 && active0) ? (int)((a.skeys[gs0 - 1] - cb) / (uint32_t)nx) : -1;
    const bool pstart = active0 && (tid == 0 || prow != hrow0);
    const int nrow_next = (tid + 1 < nact) ? (int)((a.skeys[gs0 + 1] - cb) / (uint32_t)nx) : -1;
    const bool pend = active0 && (tid + 1 == nact || nrow_next != hrow0);
    const uint64_t wstarts = __ballot(pstart);
    if (lane == 0) s_cnt[wv] = (uint32_t)__popcll(wstarts);
    __syncthreads();
    uint32_t before = 0, nparts = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        before += w < wv ? s_cnt[w] : 0u;
        nparts += s_cnt[w];
    }
    const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const int mypart0 = (int)(before + __popcll(wstarts & le)) - 1;
    if (pstart && mypart0 < kMaxParts) {
        s_part[mypart0][0] = hrow0;
        s_part[mypart0][1] = hcx0;
    }
    if (pend && mypart0 < kMaxParts) s_part[mypart0][2] = hcx0;
    __syncthreads();

    int p_row[kMaxParts], p_ca[kMaxParts], p_w[kMaxParts], p_base[kMaxParts + 1];
    p_base[0] = 0;
#pragma unroll
    for (int p = 0; p < kMaxParts; ++p) {
        const bool have = p < (int)nparts;
        p_row[p] = have ? s_part[p][0] : 0;
        p_ca[p] = have ? s_part[p][1] : 0;
        p_w[p] = have ? s_part[p][2] - p_ca[p] + 3 : 0;     // span + 2 staged columns
        p_base[p + 1] = p_base[p] + NR * p_w[p];
    }
    const int NC = p_base[kMaxParts];
    const bool use_lds = !w.all_global && a.cap > 0 && nparts <= (uint32_t)kMaxParts && NC <= a.max_cells;

    // staged cells: start and count, then an exclusive prefix in LDS (carry ends
    // as the staged particle count, the same in every thread)
    uint32_t carry = 0;
    if (use_lds) {
        for (int f0 = 0; f0 < NC; f0 += BS) {
            const int f = f0 + tid;
            uint32_t c = 0;
            if (f < NC) {
                int pb = p_base[0], pw = p_w[0], pr = p_row[0], pc = p_ca[0];
#pragma unroll
                for (int q = 1; q < kMaxParts; ++q)
                    if (f >= p_base[q] && p_w[q] > 0) { pb = p_base[q]; pw = p_w[q]; pr = p_row[q]; pc = p_ca[q]; }
                const int r = (f - pb) / pw, j = (f - pb) - r * pw;
                const int oy = (r % 3) - 1, oz = DIM == 3 ? (r / 3) - 1 : 0;
                const int cy = pr % nx, cz = DIM == 3 ? pr / nx : 0;
                const int x = wrap_cell(pc - 1 + j, nx);
                const int y = wrap_cell(cy + oy, nx);
                const int z = DIM == 3 ? wrap_cell(cz + oz, nx) : 0;
                const uint32_t ci = cb + (uint32_t)(x + y * nx + z * nx * nx);
                const uint32_t s0 = a.cell_start[ci];
                c = a.cell_end[ci] - s0;
                gs[f] = s0;
            }
            const uint32_t inc = wave_incl_scan(c);
            if (lane == kWave - 1) s_cnt[wv] = inc;
            __syncthreads();
            uint32_t wbase0 = carry, tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                wbase0 += w < wv ? s_cnt[w] : 0u;
                tot += s_cnt[w];
            }
            if (f < NC) off[f] = wbase0 + inc - c;
            carry += tot;
            __syncthreads();
        }
        if (tid == 0) off[NC] = carry;
    }
    __syncthreads();

    // Balanced path: the whole staged neighbourhood fits one window and the
    // homes' candidate runs fit the pool of mask words (a run of len
    // candidates takes ceil(len / 32) words; known here from the staged
    // offsets, so each home gets its own run of the pool).  Homes are then
    // screened in their sorted order and stepped, after a deal by their
    // kept-candidate counts, by lanes of nearly equal work.
    const int S = (int)carry;
    bool bal = false;
    if (use_lds && a.wpool > 0 && S <= a.cap) {
        int need = 0;
        if (active0) {
            int pb = p_base[0], pw = p_w[0], pc = p_ca[0];
#pragma unroll
            for (int q = 1; q < kMaxParts; ++q)
                if (mypart0 == q) { pb = p_base[q]; pw = p_w[q]; pc = p_ca[q]; }
            const int jh = hcx0 - pc + 1;
            for (int r = 0; r < NR; ++r) {
                const int rb = pb + r * pw;
                need += ((int)(off[rb + jh + 2] - off[rb + jh - 1]) + 31) >> 5;
            }
        }
        uint32_t total = 0;
        const uint32_t woff = block_excl_scan<BS>((uint32_t)need, s_cnt, total);
        s_wo[tid] = (uint16_t)woff;
        bal = total <= (uint32_t)a.wpool;
    }

    // Windowed path: deal the homes to threads grouped by where they sit in
    // their cell (y quartile in 2D, y/z halves in 3D; stable within a group).
    // A home's accepted candidates bunch in the stencil rows nearest to it, so
    // lanes of one wave then fill their candidate lists at similar rates and
    // list flushes keep more lanes busy.  Each home's own work is unchanged.
    // The balanced path screens home tid on thread tid.
    uint32_t dealt = (uint32_t)tid | ((uint32_t)max(0, min(mypart0, 255)) << 16);
    if (!bal) {
        const int dkey = active0 ? deal_key<DIM>(hy0, hz0, hrow0, nx, w.cs) : kDealKeys - 1;
        const uint64_t lt = (1ull << lane) - 1;
        uint32_t myrank = 0;
#pragma unroll
        for (int q = 0; q < kDealKeys; ++q) {
            const uint64_t b = __ballot(dkey == q);
            if (lane == 0) s_qc[wv][q] = (uint32_t)__popcll(b);
            if (dkey == q) myrank = (uint32_t)__popcll(b & lt);
        }
        __syncthreads();
        for (int q = 0; q < kDealKeys; ++q)
#pragma unroll
            for (int v = 0; v < NW; ++v)
                myrank += (q < dkey || (q == dkey && v < wv)) ? s_qc[v][q] : 0u;
        s_deal[myrank] = dealt;
        __syncthreads();
        dealt = s_deal[tid];
    }

    // this thread's home (after the deal)
    const int th = (int)(dealt & 0xFFFFu);
    const int mypart = (int)(dealt >> 16);
    const bool active = th < nact;
    const int g = g0 + (active ? th : nact - 1);
    // (the balanced path keeps home tid: its data is already here)
    const uint32_t hkey = bal ? hkey0 : a.skeys[g] - cb;
    const int hrow = (int)(hkey / (uint32_t)nx);
    const int hcx = (int)(hkey - (uint32_t)hrow * (uint32_t)nx);
    const uint32_t ho = bal ? ho0 : a.perm[g];
    const uint32_t hx = bal ? hx0 : a.sx[ho], hy = bal ? hy0 : a.sy[ho];
    const uint32_t hz = DIM == 3 ? (bal ? hz0 : a.sz[ho]) : 0u;
    const int htype = bal ? htype0 : a.type[ho];
    const int mrow = htype * w.ntypes;

    // 3D phase A screens on 16-bit positions relative to the workgroup's first
    // home: rel = (int32)(p - origin) >> sh, with sh chosen so that every
    // staged particle and home lies within 2^14 units of the origin (periodic
    // distance, in cells, bounded from the parts' staged ranges).  Relative
    // differences equal the minimum-image differences only while everything
    // lies within a quarter box of the origin (two such points are less than
    // half a box apart); wider workgroups (coarse grids, sparse parts) screen
    // on the exact 32-bit differences instead.
    int sh = 0, bq = 0;
    bool narrow = false;
    {
        const uint32_t okey = a.skeys[g0] - cb;
        const int orow = (int)(okey / (uint32_t)nx);
        const int ocx = (int)(okey - (uint32_t)orow * (uint32_t)nx);
        const int ocy = orow % nx, ocz = DIM == 3 ? orow / nx : 0;
        int ext = 0;
#pragma unroll
        for (int p = 0; p < kMaxParts; ++p) {
            if (p >= (int)nparts) break;
            const int cy = p_row[p] % nx, cz = DIM == 3 ? p_row[p] / nx : 0;
            ext = max(ext, periodic_cells(p_ca[p] - 1, ocx, nx) + p_w[p]);
            ext = max(ext, periodic_cells(cy, ocy, nx) + 2);
            if (DIM == 3) ext = max(ext, periodic_cells(cz, ocz, nx) + 2);
        }
        // + 3 cells: positions inside the particle's and the origin's cells
        // (the last cell of a row can be up to twice as wide)
        const double span_units = (double)(ext + 3) * w.cs * kFix;
        narrow = span_units < 0.25 * kFix;
        while (sh < 31 && span_units >= 16384.0 * (double)(1u << sh)) ++sh;
        // |D|^2 <= rmax^2 implies sum (D_i / 2^sh)^2 in quantised units <= (Rq + 2)^2
        const double rq = w.cs * kFix * (1.0 + 0x1p-20) / (double)(1u << sh) + 2.0;
        bq = (int)fmin(2147483647.0, floor(rq * rq) + 1.0);
    }

    // The force law's two additive constants held in VGPRs: each meets
    // another scalar operand in an FMA, which reads one scalar at most.
    WorldDev wf = w;
    asm volatile("" : "+v"(wf.one_plus_beta));
    asm volatile("" : "+v"(wf.inv_beta));
    // One candidate k of the staged window for a home at (px, py, pz) with
    // matrix row prow_m: the reference's FP64 predicate and force law.
    auto visit = [&](Acc &ac, int k, uint32_t px, uint32_t py, uint32_t pz, int prow_m) {
        const uint2 q = sxy[k];
        const int dxi = (int)(q.x - px), dyi = (int)(q.y - py);
        const int dzi = DIM == 3 ? (int)(szz[k] - pz) : 0;
        bool ok = true;
        if constexpr (!WRAP) {
            // plain difference of the clamped world: |D| >= 1/2 unless the
            // int32 difference has the sign of the unsigned order
            ok = ((dxi >= 0) == (q.x >= px)) & ((dyi >= 0) == (q.y >= py));
            if (DIM == 3) ok &= ((dzi >= 0) == (szz[k] >= pz));
        }
        const int idx = prow_m + (int)st[k];
        const double m = MLDS ? M[idx] : a.matrix[w.moff + idx] * w.R;
        pair64_sel<DIM, DEBUG>(ac, ok, dxi, dyi, dzi, m, DEBUG ? sid[k] : 0u, wf);
    };
    // Stages window [w0, w0 + wn) of the staged neighbourhood.
    auto stage = [&](int w0, int wn) {
        for (int s = tid; s < wn; s += BS) {
            const int sg = w0 + s;
            int lo = 0, hi = NC;           // off[lo] <= sg < off[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (off[mid] <= (uint32_t)sg) lo = mid;
                else hi = mid;
            }
            const uint32_t q = a.perm[gs[lo] + ((uint32_t)sg - off[lo])];
            const uint32_t qx = a.sx[q], qy = a.sy[q];
            sxy[s] = make_uint2(qx, qy);
            if constexpr (DIM == 3) {
                const uint32_t qz = a.sz[q];
                szz[s] = qz;
                sq[s] = rel16(qx, ox, sh) | (rel16(qy, oy, sh) << 16);
                sqz[s] = (uint16_t)rel16(qz, oz, sh);
            }
            if constexpr (DEBUG) sid[s] = (uint32_t)a.id[q];
            st[s] = (uint8_t)a.type[q];
        }
    };

    Acc acc;
    acc.x = acc.y = acc.z = 0.0;
    acc.cnt = 0;
    acc.hash = 0ull;
    unsigned long long mycnt = 0;      // interacting pairs of the homes this thread finished
    if (use_lds) {
        int pb = p_base[0], pw = p_w[0], pc = p_ca[0];
#pragma unroll
        for (int q = 1; q < kMaxParts; ++q)
            if (mypart == q) { pb = p_base[q]; pw = p_w[q]; pc = p_ca[q]; }
        const int jh = hcx - pc + 1;           // staged column of the home cell
        // Reference quirk: the stencil origin floor(x/cs) may be nx while the
        // key is clamped to nx-1 (Physics.java:367-372 vs 404-405).  The
        // reference then scans columns {nx-1, 0, 1}: skip staged column nx-2
        // (column 1 is always farther than rmax).  Same for rows / layers.
        const int cy = hrow % nx, cz = DIM == 3 ? hrow / nx : 0;
        // (the FP64 division only for homes in a last column, row or layer)
        bool q2x = false, q2y = false, q2z = false;
        if (hcx == nx - 1) q2x = stencil_origin(hx, w.cs) == nx;
        if (cy == nx - 1) q2y = stencil_origin(hy, w.cs) == nx;
        if (DIM == 3 && cz == nx - 1) q2z = stencil_origin(hz, w.cs) == nx;
        const int j_lo = jh - 1 + (q2x ? 1 : 0);
        const uint32_t hq = rel16(hx, ox, sh) | (rel16(hy, oy, sh) << 16);
        const uint32_t hqz = DIM == 3 ? rel16(hz, oz, sh) : 0u;

        // U candidates from base kb of a row run of len (lane-local offset i):
        // bit u set when candidate kb + u passes the phase-A screen
        // NARROW (3D): on the staged 16-bit copies; otherwise FP32 on the
        // exact differences.  Bits at or past the run's end (i + u >= len)
        // stay clear.
        auto screen_group = [&](auto narrow_tag, int kb, int i, int len) -> uint32_t {
            constexpr bool NARROW = decltype(narrow_tag)::value;
            uint32_t keep = 0;
            if constexpr (NARROW && DIM == 3) {
                const uint32_t *qp = sq + kb;
                const uint16_t *qzp = sqz + kb;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    keep |= screen16<DIM>(qp[u], hq, (uint32_t)qzp[u], hqz, bq) ? (1u << u) : 0u;
            } else {
                const uint2 *qp = sxy + kb;
                const uint32_t *qzp = szz + kb;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    keep |= screen<DIM>(hx, hy, hz, qp[u].x, qp[u].y, DIM == 3 ? qzp[u] : 0u, w.r2hi)
                                ? (1u << u) : 0u;
            }
            const int nv = min(max(len - i, 0), U);
            return keep & ((1u << nv) - 1u);
        };

        if (bal) {
            // ---- balanced path: one window -------------------------------
            stage(0, S);
            __syncthreads();
            // Phase A: thread tid screens home tid; each 32 candidates of a row
            // run give one mask word, stored (with its first candidate's index)
            // when any candidate passed.
            int nwd = 0, npass = 0;
            const int wo = s_wo[tid];          // this home's run of the word pool
            auto scan_words = [&](auto narrow_tag) {
#pragma unroll 1
                for (int r = 0; r < NR; ++r) {
                    const int oyr = (r % 3) - 1, ozr = DIM == 3 ? (r / 3) - 1 : 0;
                    const bool skip = !active || (q2y && oyr == -1) || (q2z && ozr == -1);
                    const int rb = pb + r * pw;
                    const int klo = (int)off[rb + j_lo];
                    const int len = skip ? 0 : max(0, (int)off[rb + jh + 2] - klo);
                    const int K = wave_max(len);
                    for (int i = 0; i < K; i += 4 * U) {
                        uint32_t word = 0;
#pragma unroll
                        for (int gq = 0; gq < 4; ++gq) {
                            const int ig = i + gq * U;
                            if (ig < K) {
                                // a lane past its run reads the window's first entries
                                // (and keeps none of them)
                                const int kb = ig < len ? klo + ig : 0;
                                word |= screen_group(narrow_tag, kb, ig, len) << (gq * U);
                            }
                        }
                        if (word) {
                            wmask[wo + nwd] = word;
                            wbase[wo + nwd] = (uint16_t)(klo + i);
                            ++nwd;
                            npass += __popc(word);
                        }
                    }
                }
            };
            // 2D screens on the 32-bit differences: its FP32 screen is short;
            // 3D gains 2/3 of its screen on the 16-bit copies
            if constexpr (DIM == 3) {
                if (narrow) scan_words(std::true_type{});
                else scan_words(std::false_type{});
            } else {
                scan_words(std::false_type{});
            }
            int si = 0;
            if (active) {
                const int f = pb + RMID * pw + jh;
                si = (int)off[f] + (g - (int)gs[f]);
            }
            s_si[tid] = (uint16_t)si;
            s_hc[tid] = (uint16_t)(active ? npass : 0);
            // Deal the homes by their kept-candidate counts (a counting sort;
            // the order inside a count does not matter): lane tid of the
            // workgroup steps the home of rank tid, so the lanes of a wave
            // carry nearly equal work.  Each home is stepped whole by one lane
            // in its scan order, so its sums are those of the windowed path.
            for (int t = tid; t < 256; t += BS) s_bin[t] = 0u;
            __syncthreads();
            const int cbin = min(active ? npass : 0, 255);
            const uint32_t local = atomicAdd(&s_bin[cbin], 1u);
            __syncthreads();
            {
                uint32_t v[PER], sum = 0, total = 0;
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    v[q] = s_bin[tid * PER + q];
                    sum += v[q];
                }
                uint32_t ex = block_excl_scan<BS>(sum, s_cnt, total);
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    s_bin[tid * PER + q] = ex;
                    ex += v[q];
                }
            }
            __syncthreads();
            s_deal[s_bin[cbin] + local] = (uint32_t)tid;
            __syncthreads();

            // Phase B: the home of rank tid, its kept candidates word by word
            const int h = (int)s_deal[tid];
            if (h < nact) {
                const int hsi = s_si[h];
                const uint2 hp = sxy[hsi];
                const uint32_t bhz = DIM == 3 ? szz[hsi] : 0u;
                const int btype = (int)st[hsi];
                const int bmrow = btype * w.ntypes;
                Acc ab;
                ab.x = ab.y = ab.z = 0.0;
                ab.cnt = 0;
                ab.hash = 0ull;
                int left = s_hc[h], j = s_wo[h], kb = 0;
                uint32_t m = 0;
                auto next = [&]() {
                    if (m == 0) {
                        m = wmask[j];
                        kb = wbase[j];
                        ++j;
                    }
                    const int k = kb + __builtin_ctz(m);
                    m &= m - 1u;
                    return k;
                };
                // two candidates per iteration (half the loop's bookkeeping;
                // their force laws overlap), added in candidate order
#pragma unroll 1
                for (; left >= 2; left -= 2) {
                    const int k1 = next();
                    const int k2 = next();
                    visit(ab, k1, hp.x, hp.y, bhz, bmrow);
                    visit(ab, k2, hp.x, hp.y, bhz, bmrow);
                }
                if (left > 0) visit(ab, next(), hp.x, hp.y, bhz, bmrow);
                const int gh = g0 + h;
                finish_home<DIM, WRAP, DEBUG>(a, w, ab, hp.x, hp.y, bhz, btype, a.perm[gh], gh);
                mycnt = (unsigned long long)ab.cnt;
            }
        } else {
            // ---- windowed path --------------------------------------------
            // The staged neighbourhood is S particles; it is brought into LDS in
            // windows of at most cap (one window for ordinary densities, several
            // for clustered cells) and every lane scans its ranges clipped to it.
            for (int w0 = 0; w0 < S; w0 += a.cap) {
                const int wn = min(a.cap, S - w0);
                stage(w0, wn);
                __syncthreads();

                // Per-lane candidate list: entries at lists[j * BS + tid]; lp walks
                // the lane's column of the list in steps of BS entries.
                uint16_t *const lbase = lists + tid;
                uint16_t *lp = lbase;
                uint16_t *const lfull = lbase + (a.lcap - U) * BS;     // flush beyond this
                auto flush = [&]() {
                    for (uint16_t *e = lbase; e < lp; e += BS) visit(acc, *e, hx, hy, hz, mrow);
                    lp = lbase;
                };
                // the row scan, instantiated for the 16-bit and the 32-bit screen
                auto scan_rows = [&](auto narrow_tag) {
#pragma unroll 1
                    for (int r = 0; r < NR; ++r) {
                        const int oyr = (r % 3) - 1, ozr = DIM == 3 ? (r / 3) - 1 : 0;
                        const bool skip = !active || (q2y && oyr == -1) || (q2z && ozr == -1);
                        const int rb = pb + r * pw;
                        const int klo = max((int)off[rb + j_lo], w0) - w0;
                        const int khi = min((int)off[rb + jh + 2], w0 + wn) - w0;
                        const int len = skip ? 0 : max(0, khi - klo);
                        const int K = wave_max(len);
                        for (int i = 0; i < K; i += U) {
                            if (__any(lp > lfull)) flush();
                            // U candidates from one base: a lane past its range reads the
                            // window's first entries instead (and keeps none of them)
                            const int kb = i < len ? klo + i : 0;
                            const uint32_t keep = screen_group(narrow_tag, kb, i, len);
                            // branch-free append: every candidate is written to the next
                            // free slot, which only a kept one claims (the flush check
                            // above leaves room for U slots)
#pragma unroll
                            for (int u = 0; u < U; ++u) {
                                *lp = (uint16_t)(kb + u);
                                lp += ((keep >> u) & 1u) ? BS : 0;
                            }
                        }
                    }
                };
                if constexpr (DIM == 3) {
                    if (narrow) scan_rows(std::true_type{});
                    else scan_rows(std::false_type{});
                } else {
                    scan_rows(std::false_type{});
                }
                flush();
                __syncthreads();
            }
        }
    } else if (active) {
        home_global<DIM, WRAP, DEBUG, MLDS>(a, w, M, g, hx, hy, hz, mrow, acc);
    }
    if (!bal && active) {
        finish_home<DIM, WRAP, DEBUG>(a, w, acc, hx, hy, hz, htype, ho, g);
        mycnt = (unsigned long long)acc.cnt;
    }

    // interacting pairs, one 64-bit atomic per workgroup
    unsigned long long tot = mycnt;
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) tot += __shfl_xor(tot, d, kWave);
    if (lane == 0) s_pairs[wv] = tot;
    __syncthreads();
    if (tid == 0) {
        unsigned long long sum = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) sum += s_pairs[w];
        if (sum) {
            atomicAdd(a.pair_total, sum);
            atomicAdd(a.pair_world + bw.x, sum);
        }
        record_block_work(a, blk, sum);
    }
}

// Worlds whose workgroups would span more grid rows than the staged layout
// holds (a particle or less per cell): one thread per home of the cell-sorted
// order, each running home_global, with no LDS but the type matrix, so that
// many waves per CU hide the latency of the cell bounds and the gathers.
// Results are those of k_force for the same homes, bit for bit (k_force's
// global path is this code, and its staged path visits the same candidates in
// the same order).
constexpr int kSparseBlock = 256;

template <int DIM, bool WRAP, bool DEBUG, bool MLDS>
__global__ __launch_bounds__(kSparseBlock) void k_force_sparse(ForceArgs a)
{
    constexpr int NW = kSparseBlock / kWave;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned long long s_pairs[NW];
    double *M = reinterpret_cast<double *>(smem);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const uint2 bw = a.blocks[a.block_base + blk];
    const WorldDev w = a.worlds[bw.x];
    if constexpr (MLDS) {
        for (int t = tid; t < w.ntypes * w.ntypes; t += kSparseBlock) M[t] = a.matrix[w.moff + t] * w.R;
        __syncthreads();
    }
    const int g = (int)bw.y + tid;
    const bool active = g < w.off + w.n;
    Acc acc;
    acc.x = acc.y = acc.z = 0.0;
    acc.cnt = 0;
    acc.hash = 0ull;
    if (active) {
        const uint32_t ho = a.perm[g];
        const uint32_t hx = a.sx[ho], hy = a.sy[ho], hz = DIM == 3 ? a.sz[ho] : 0u;
        const int htype = a.type[ho];
        home_global<DIM, WRAP, DEBUG, MLDS>(a, w, M, g, hx, hy, hz, htype * w.ntypes, acc);
        finish_home<DIM, WRAP, DEBUG>(a, w, acc, hx, hy, hz, htype, ho, g);
    }
    // interacting pairs, one 64-bit atomic per workgroup
    unsigned long long tot = active ? (unsigned long long)acc.cnt : 0ull;
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) tot += __shfl_xor(tot, d, kWave);
    if (lane == 0) s_pairs[wv] = tot;
    __syncthreads();
    if (tid == 0) {
        unsigned long long sum = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) sum += s_pairs[k];
        if (sum) {
            atomicAdd(a.pair_total, sum);
            atomicAdd(a.pair_world + bw.x, sum);
        }
        record_block_work(a, blk, sum);
    }
}

}  // namespace plife
