// plife_force.hpp -- the fused velocity + position kernel of the particle-life step.
//
// Reference: backend/Physics.java:397-450 (updateVelocity, updatePosition),
// the accelerator lambda app/Main.java:275-280, Range.java:40-96.
//
// Work decomposition: one 256-thread workgroup per strip of W cells along x in
// one grid row (2D: row = cy, 3D: row = cy + cz*nx).  The workgroup stages the
// 3^(DIM-1) neighbour rows x (W+2) cells of the cell-sorted SoA in LDS; each
// thread then owns one home particle of the strip and, per stencil row, scans
// the three staged cells around its own column -- exactly the cells the
// reference's 3x3 (3x3x3) stencil visits.
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

#include "plife_kernels.hpp"

namespace plife {

struct ForceArgs {
    // cell-sorted step-start state
    const uint32_t *sx, *sy, *sz;
    float *vx, *vy, *vz;             // updated in place (each particle writes its own)
    const int32_t *type, *id;
    // new positions, same (sorted) order
    uint32_t *ox, *oy, *oz;
    const uint32_t *cell_start, *cell_end;
    const double *matrix;            // ntypes*ntypes, row = acted-on type
    int ntypes, m_in_lds;
    int n, nx, W, nstrips, cap, lcap, all_global;
    float r2hi;                      // screen bound, units 2^-64: (rmax 2^32)^2 (1 + 2^-20)
    double r2u;                      // fl(rmax*rmax) * 2^64: the reference's bound in units 2^-64
    double S;                        // 2^-32 / rmax: |rel| = |D| S with D in units 2^-32
    double R;                        // rmax * 2^32:  1/|rel| = R / |D|
    double beta, one_plus_beta, inv_1mb, inv_beta;
    double ff;                       // friction^(60 dt) (host FP64 pow, Physics.java:401)
    double kdt;                      // force * dt * 2^-32
    double cs, dt;
    unsigned long long *pair_counter;
    int32_t *dbg_count;
    unsigned long long *dbg_hash;
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
template <int DIM, bool DEBUG>
__device__ __forceinline__ void pair64(Acc &acc, int dxi, int dyi, int dzi, double m, uint32_t qid,
                                       const ForceArgs &a)
{
    const double dx = (double)dxi, dy = (double)dyi, dz = (double)dzi;
    double d2;
    {
#pragma clang fp contract(off)
        // JOML lengthSquared without FMA: x*x + (y*y + z*z) (exact 2^64 scaling)
        d2 = DIM == 3 ? dx * dx + (dy * dy + dz * dz) : dx * dx + dy * dy;
    }
    const bool in = (d2 != 0.0) & (d2 <= a.r2u);     // Physics.java:432
    const double y0 = (double)__builtin_amdgcn_rsqf((float)d2);
    const double e = d2 * y0;
    const double h = fma(-e, y0, 1.0);
    const double y = fma(0.5 * y0, h, y0);            // 1/|D|, one Newton step from FP32
    const double dist = (d2 * y) * a.S;
    const double inv = y * a.R;                        // 1/dist
    const double t = fabs(fma(dist, -2.0, a.one_plus_beta));
    const double g = fma(t, -a.inv_1mb, 1.0);
    const double att = (m * g) * inv;
    const double rep = a.inv_beta - inv;
    double c = dist < a.beta ? rep : att;
    c = in ? c : 0.0;
    acc.x = fma(c, dx, acc.x);
    acc.y = fma(c, dy, acc.y);
    if constexpr (DIM == 3) acc.z = fma(c, dz, acc.z);
    acc.cnt += in ? 1 : 0;
    if constexpr (DEBUG) {
        if (in) acc.hash += id_hash(qid);
    }
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, kWave));
    return v;
}

// Phase A screen on exact fixed-point deltas.  In a clamped world the int32
// difference equals the plain difference only when its sign matches the
// unsigned order; otherwise |D| >= 1/2 > rmax and the pair is dropped.
template <int DIM, bool WRAP>
__device__ __forceinline__ bool screen(uint32_t hx, uint32_t hy, uint32_t hz, uint32_t qx,
                                       uint32_t qy, uint32_t qz, float r2hi)
{
    const int dxi = (int)(qx - hx), dyi = (int)(qy - hy);
    const float dx = (float)dxi, dy = (float)dyi;
    float d2 = dx * dx;
    d2 = fmaf(dy, dy, d2);
    int dzi = 0;
    if constexpr (DIM == 3) {
        dzi = (int)(qz - hz);
        const float dz = (float)dzi;
        d2 = fmaf(dz, dz, d2);
    }
    bool pass = (d2 > 0.0f) & (d2 <= r2hi);
    if constexpr (!WRAP) {
        pass &= ((dxi >= 0) == (qx >= hx)) & ((dyi >= 0) == (qy >= hy));
        if constexpr (DIM == 3) pass &= ((dzi >= 0) == (qz >= hz));
    }
    return pass;
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
__device__ __forceinline__ void finish_home(const ForceArgs &a, const Acc &acc, uint32_t hx,
                                            uint32_t hy, uint32_t hz, int g)
{
    const double nvx = a.ff * (double)a.vx[g] + a.kdt * acc.x;
    const double nvy = a.ff * (double)a.vy[g] + a.kdt * acc.y;
    a.vx[g] = (float)nvx;
    a.vy[g] = (float)nvy;
    a.ox[g] = integrate_axis(hx, nvx, a.dt, WRAP);
    a.oy[g] = integrate_axis(hy, nvy, a.dt, WRAP);
    if constexpr (DIM == 3) {
        const double nvz = a.ff * (double)a.vz[g] + a.kdt * acc.z;
        a.vz[g] = (float)nvz;
        a.oz[g] = integrate_axis(hz, nvz, a.dt, WRAP);
    }
    if constexpr (DEBUG) {
        a.dbg_count[g] = acc.cnt;
        a.dbg_hash[g] = acc.hash;
    }
}

// Literal reference stencil over 3^DIM cells straight from global memory
// (Physics.java:404-424), FP64 throughout.  Used for strips whose staged
// neighbourhood does not fit in LDS and for grids with fewer than 3 cells per
// axis (where the reference visits repeated cells).
template <int DIM, bool WRAP, bool DEBUG>
__device__ void home_global(const ForceArgs &a, const double *Mlds, int g, uint32_t hx,
                            uint32_t hy, uint32_t hz, int row, Acc &acc)
{
    const int nx = a.nx;
    const int cx0 = stencil_origin(hx, a.cs);
    const int cy0 = stencil_origin(hy, a.cs);
    const int cz0 = DIM == 3 ? stencil_origin(hz, a.cs) : 0;
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
                const int ci = cx + cy * nx + cz * nx * nx;
                const int lo = (int)a.cell_start[ci], hi = (int)a.cell_end[ci];
                for (int j = lo; j < hi; ++j) {
                    if (j == g) continue;
                    const int dxi = (int)(a.sx[j] - hx), dyi = (int)(a.sy[j] - hy);
                    const int dzi = DIM == 3 ? (int)(a.sz[j] - hz) : 0;
                    if (!WRAP) {
                        // plain difference of the clamped world: skip |D| >= 1/2
                        bool ok = ((dxi >= 0) == (a.sx[j] >= hx)) & ((dyi >= 0) == (a.sy[j] >= hy));
                        if (DIM == 3) ok &= ((dzi >= 0) == (a.sz[j] >= hz));
                        if (!ok) continue;
                    }
                    const int idx = row + a.type[j];
                    const double m = a.m_in_lds ? Mlds[idx] : a.matrix[idx];
                    pair64<DIM, DEBUG>(acc, dxi, dyi, dzi, m, DEBUG ? (uint32_t)a.id[j] : 0u, a);
                }
            }
}

// LDS layout of k_force, shared with the host's size computation.
struct ForceLds {
    size_t m, sxy, lists, sz, sid, off, gs, st, total;
};

__host__ __device__ inline ForceLds force_lds_layout(int dim, bool debug, int ntypes, bool m_in_lds,
                                                     int cap, int lcap, int W)
{
    const int NR = dim == 3 ? 9 : 3;
    ForceLds L;
    size_t o = 0;
    L.m = o;     o += m_in_lds ? (size_t)ntypes * ntypes * 8 : 0;
    L.sxy = o;   o += (size_t)cap * 8;
    L.lists = o; o += (size_t)lcap * kBlock * 2;
    o = (o + 3) & ~(size_t)3;
    L.sz = o;    o += dim == 3 ? (size_t)cap * 4 : 0;
    L.sid = o;   o += debug ? (size_t)cap * 4 : 0;
    L.off = o;   o += (size_t)(NR * (W + 2) + 1) * 4;
    L.gs = o;    o += (size_t)(NR * (W + 2)) * 4;
    L.st = o;    o += (size_t)cap;
    L.total = (o + 15) & ~(size_t)15;
    return L;
}

template <int DIM, bool WRAP, bool DEBUG>
__global__ __launch_bounds__(kBlock) void k_force(ForceArgs a)
{
    constexpr int NR = DIM == 3 ? 9 : 3;      // stencil rows
    constexpr int RC = DIM == 3 ? 4 : 1;      // centre row (oy = oz = 0)
    constexpr int U = 4;                      // screen unroll; lists keep U slots free
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t s_w[4];
    __shared__ unsigned long long s_w64[4];

    const int nx = a.nx;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int strip = b % a.nstrips, row = b / a.nstrips;
    const int cy = row % nx, cz = DIM == 3 ? row / nx : 0;
    const int cx0 = strip * a.W;
    const int wlen = min(a.W, nx - cx0);
    const int rc = wlen + 2;                  // staged cells per row
    const int NC = NR * rc;

    const ForceLds L = force_lds_layout(DIM, DEBUG, a.ntypes, a.m_in_lds != 0, a.cap, a.lcap, a.W);
    double *M = reinterpret_cast<double *>(smem + L.m);
    uint2 *sxy = reinterpret_cast<uint2 *>(smem + L.sxy);
    uint16_t *lists = reinterpret_cast<uint16_t *>(smem + L.lists);
    uint32_t *szz = reinterpret_cast<uint32_t *>(smem + L.sz);
    uint32_t *sid = reinterpret_cast<uint32_t *>(smem + L.sid);
    uint32_t *off = reinterpret_cast<uint32_t *>(smem + L.off);
    uint32_t *gs = reinterpret_cast<uint32_t *>(smem + L.gs);
    uint8_t *st = reinterpret_cast<uint8_t *>(smem + L.st);

    const int tid = threadIdx.x;
    if (a.m_in_lds)
        for (int t = tid; t < a.ntypes * a.ntypes; t += kBlock) M[t] = a.matrix[t];

    // per staged cell: global start and count, then a flat exclusive prefix
    uint32_t cnt3[3];
    uint32_t csum = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int f = tid * 3 + k;
        uint32_t c = 0;
        if (f < NC) {
            const int r = f / rc, col = f - r * rc;
            const int oy = (r % 3) - 1, oz = DIM == 3 ? (r / 3) - 1 : 0;
            const int x = wrap_cell(cx0 - 1 + col, nx);
            const int y = wrap_cell(cy + oy, nx);
            const int z = DIM == 3 ? wrap_cell(cz + oz, nx) : 0;
            const int ci = x + y * nx + z * nx * nx;
            const uint32_t s0 = a.cell_start[ci];
            c = a.cell_end[ci] - s0;
            gs[f] = s0;
        }
        cnt3[k] = c;
        csum += c;
    }
    uint32_t total;
    uint32_t o = block_excl_scan(csum, s_w, total);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int f = tid * 3 + k;
        if (f < NC) off[f] = o;
        o += cnt3[k];
    }
    if (tid == 0) off[NC] = total;
    __syncthreads();

    const int S = (int)total;
    const int fh = RC * rc + 1;                // first home cell (flat index)
    const int nhome = (int)(off[fh + wlen] - off[fh]);
    const int ghome = (int)gs[fh];             // home cells are contiguous in global memory
    const bool use_lds = !a.all_global && S <= a.cap;

    if (use_lds) {
        for (int s = tid; s < S; s += kBlock) {
            int lo = 0, hi = NC;               // off[lo] <= s < off[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (off[mid] <= (uint32_t)s) lo = mid;
                else hi = mid;
            }
            const int g = (int)(gs[lo] + ((uint32_t)s - off[lo]));
            sxy[s] = make_uint2(a.sx[g], a.sy[g]);
            if constexpr (DIM == 3) szz[s] = a.sz[g];
            if constexpr (DEBUG) sid[s] = (uint32_t)a.id[g];
            st[s] = (uint8_t)a.type[g];
        }
    }
    __syncthreads();

    long long my_pairs = 0;
    // all lanes iterate together (inactive lanes carry empty ranges) so the
    // wave-uniform candidate loops and list flushes below see every lane
    for (int base = 0; base < nhome; base += kBlock) {
        const int hh = base + tid;
        const bool active = hh < nhome;
        const int g = ghome + (active ? hh : 0);
        Acc acc;
        acc.x = acc.y = acc.z = 0.0;
        acc.cnt = 0;
        acc.hash = 0ull;
        uint32_t hx = 0, hy = 0, hz = 0;
        int hrow = 0;
        if (use_lds) {
            const int hl = (int)off[fh] + (active ? hh : 0);
            const uint2 e = sxy[hl];
            hx = e.x;
            hy = e.y;
            if constexpr (DIM == 3) hz = szz[hl];
            hrow = (int)st[hl] * a.ntypes;
            // home column c in [1, wlen]: off[RC*rc + c] <= hl < off[RC*rc + c + 1]
            int lo = 1, hi = wlen + 1;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (off[RC * rc + mid] <= (uint32_t)hl) lo = mid;
                else hi = mid;
            }
            const int c = lo;
            // Reference quirk: the stencil origin floor(x/cs) may be nx while the
            // key is clamped to nx-1 (Physics.java:367-372 vs 404-405).  The
            // reference then scans columns {nx-1, 0, 1}: skip staged column nx-2
            // (column 1 is always farther than rmax).  Same for rows / layers.
            const bool q2x = (cx0 + c - 1 == nx - 1) && stencil_origin(hx, a.cs) == nx;
            const bool q2y = (cy == nx - 1) && stencil_origin(hy, a.cs) == nx;
            const bool q2z = DIM == 3 && (cz == nx - 1) && stencil_origin(hz, a.cs) == nx;
            const int c_lo = c - 1 + (q2x ? 1 : 0);

            int cnt = 0;
            auto flush = [&]() {
                for (int j = 0; j < cnt; ++j) {
                    const int k = lists[j * kBlock + tid];
                    const uint2 q = sxy[k];
                    const int dxi = (int)(q.x - hx), dyi = (int)(q.y - hy);
                    const int dzi = DIM == 3 ? (int)(szz[k] - hz) : 0;
                    const int idx = hrow + (int)st[k];
                    const double m = a.m_in_lds ? M[idx] : a.matrix[idx];
                    pair64<DIM, DEBUG>(acc, dxi, dyi, dzi, m, DEBUG ? sid[k] : 0u, a);
                }
                cnt = 0;
            };
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                const int oy = (r % 3) - 1, oz = DIM == 3 ? (r / 3) - 1 : 0;
                const bool skip = !active || (q2y && oy == -1) || (q2z && oz == -1);
                const int klo = (int)off[r * rc + c_lo];
                const int len = skip ? 0 : (int)off[r * rc + c + 2] - klo;
                const int K = wave_max(len);
                for (int i = 0; i < K; i += U) {
                    if (__any(cnt > a.lcap - U)) flush();
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int ii = i + u;
                        if (ii < len) {
                            const int k = klo + ii;
                            const uint2 q = sxy[k];
                            const uint32_t qz = DIM == 3 ? szz[k] : 0u;
                            if (screen<DIM, WRAP>(hx, hy, hz, q.x, q.y, qz, a.r2hi)) {
                                lists[cnt * kBlock + tid] = (uint16_t)k;
                                ++cnt;
                            }
                        }
                    }
                }
            }
            flush();
        } else if (active) {
            hx = a.sx[g];
            hy = a.sy[g];
            hz = DIM == 3 ? a.sz[g] : 0u;
            hrow = a.type[g] * a.ntypes;
            home_global<DIM, WRAP, DEBUG>(a, M, g, hx, hy, hz, hrow, acc);
        }
        if (active) {
            my_pairs += acc.cnt;
            finish_home<DIM, WRAP, DEBUG>(a, acc, hx, hy, hz, g);
        }
    }

    // interacting pairs, one 64-bit atomic per workgroup
    const unsigned long long tot_pairs = block_sum_u64((unsigned long long)my_pairs, s_w64);
    if (tid == 0 && tot_pairs) atomicAdd(a.pair_counter, tot_pairs);
}

}  // namespace plife
