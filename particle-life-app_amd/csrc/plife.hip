// plife.hip -- C ABI (include/plife.h) of the gfx950 particle-life physics step.
//
// Host orchestration of one Physics.update() (backend/Physics.java:112-134):
//   bin (keys + cell histogram + scan) -> stable radix sort -> gather -> force+integrate,
// all enqueued on the handle's own stream.  No torch, no CPU fallback: every
// step runs the HIP kernels of plife_kernels.hpp and plife_force.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/plife.h"
#include "plife_force.hpp"

using namespace plife;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(PLIFE_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK()                                                                  \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return fail(PLIFE_ERR_DEVICE, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

constexpr int kMaxTypes = 256;        // the app caps the matrix size at 256 (app/Main.java:745)
constexpr int kMaxTypesLds = 32;      // matrix staged in LDS up to 32x32 (8 KB of FP64)
constexpr long long kMaxCells = 1LL << 28;
constexpr int kMaxLds = 65536;
constexpr int kClusteredLds = 17152;  // force-kernel LDS in clustered worlds: 9 workgroups per CU

struct ForcePlan {
    int bs, lcap, max_cells, cap;
    size_t lds;
};

template <typename T>
int dev_alloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PLIFE_ERR_MEMORY, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    }
    return PLIFE_OK;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace

struct plife_physics {
    int device = 0;
    hipStream_t stream = nullptr;

    // settings
    bool have_settings = false;
    plife_settings st{};
    int ntypes = 0;
    std::vector<double> matrix_host;
    double *d_matrix = nullptr;

    // particle state: positions ping-pong P[2], the rest A[2]
    int n = 0, cap = 0;
    uint32_t *px[2] = {}, *py[2] = {}, *pz[2] = {};
    float *vx[2] = {}, *vy[2] = {}, *vz[2] = {};
    int32_t *type[2] = {}, *id[2] = {};
    int pc = 0, ac = 0;

    // binning / sort scratch
    uint32_t *keys0 = nullptr, *kA = nullptr, *kB = nullptr, *vA = nullptr, *vB = nullptr;
    uint32_t *skeys = nullptr;         // cell key of each sorted slot (last radix pass)
    uint32_t *perm = nullptr;          // points at vA or vB after a step
    long long cells_cap = 0;
    uint32_t *cell_count = nullptr, *cell_start = nullptr, *cell_end = nullptr;
    size_t hist_cap = 0, bsum_cap = 0;
    uint32_t *hist = nullptr, *bsum = nullptr;
    unsigned long long *d_pairs = nullptr;   // [0] interacting pairs, [1] sum of cell counts^2
    uint64_t *stat_host = nullptr;           // pinned copy of d_pairs[1], written every step
    double occupancy = 0.0;                  // particle-weighted cell occupancy seen by the planner
    bool clustered = false;                  // the planner's regime (with hysteresis)

    // grid of the last step
    bool stepped = false;
    int last_nx = 0;
    long long last_ncells = 0;

    // debug
    int debug = 0;
    int dbg_cap = 0;
    int32_t *dbg_count = nullptr;
    unsigned long long *dbg_hash = nullptr;

    std::atomic<int> stop{0};
    bool cells_dirty = false;      // cell_count may hold stale counts

    // HIP graph of two consecutive steps (buffer roles repeat every 2 steps)
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> gkey;

    // launch tuning (0 = automatic) and what the last force launch used
    plife_tuning tune{};
    size_t last_lds = 0;
    int last_cap = 0, last_bs = 0, last_lcap = 0;

    PosBuf P(int k) { return PosBuf{px[k], py[k], pz[k]}; }
    AuxBuf A(int k) { return AuxBuf{vx[k], vy[k], vz[k], type[k], id[k]}; }
};

namespace {

int set_device(plife_physics *h)
{
    HIP_TRY(hipSetDevice(h->device));
    return PLIFE_OK;
}

void free_state(plife_physics *h)
{
    for (int k = 0; k < 2; ++k) {
        dev_free(h->px[k]);
        dev_free(h->py[k]);
        dev_free(h->pz[k]);
        dev_free(h->vx[k]);
        dev_free(h->vy[k]);
        dev_free(h->vz[k]);
        dev_free(h->type[k]);
        dev_free(h->id[k]);
    }
    dev_free(h->keys0);
    dev_free(h->kA);
    dev_free(h->kB);
    dev_free(h->vA);
    dev_free(h->vB);
    dev_free(h->skeys);
    h->perm = nullptr;
    h->cap = 0;
}

int ensure_capacity(plife_physics *h, int n)
{
    if (n <= h->cap && h->cap > 0) return PLIFE_OK;
    free_state(h);
    const int cap = std::max(n, 1);
    int rc;
    for (int k = 0; k < 2; ++k) {
        if ((rc = dev_alloc(&h->px[k], cap))) return rc;
        if ((rc = dev_alloc(&h->py[k], cap))) return rc;
        if ((rc = dev_alloc(&h->pz[k], cap))) return rc;
        if ((rc = dev_alloc(&h->vx[k], cap))) return rc;
        if ((rc = dev_alloc(&h->vy[k], cap))) return rc;
        if ((rc = dev_alloc(&h->vz[k], cap))) return rc;
        if ((rc = dev_alloc(&h->type[k], cap))) return rc;
        if ((rc = dev_alloc(&h->id[k], cap))) return rc;
    }
    if ((rc = dev_alloc(&h->keys0, cap))) return rc;
    if ((rc = dev_alloc(&h->kA, cap))) return rc;
    if ((rc = dev_alloc(&h->kB, cap))) return rc;
    if ((rc = dev_alloc(&h->vA, cap))) return rc;
    if ((rc = dev_alloc(&h->vB, cap))) return rc;
    if ((rc = dev_alloc(&h->skeys, cap))) return rc;
    h->cap = cap;
    return PLIFE_OK;
}

int ensure_cells(plife_physics *h, long long ncells)
{
    if (ncells <= h->cells_cap) return PLIFE_OK;
    dev_free(h->cell_count);
    dev_free(h->cell_start);
    dev_free(h->cell_end);
    h->cells_cap = 0;
    int rc;
    if ((rc = dev_alloc(&h->cell_count, (size_t)ncells))) return rc;
    if ((rc = dev_alloc(&h->cell_start, (size_t)ncells))) return rc;
    if ((rc = dev_alloc(&h->cell_end, (size_t)ncells))) return rc;
    // the step keeps cell_count all-zero between steps (the scan clears it)
    HIP_TRY(hipMemsetAsync(h->cell_count, 0, (size_t)ncells * 4, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->cells_cap = ncells;
    return PLIFE_OK;
}

int ensure_u32(uint32_t **p, size_t *capv, size_t need)
{
    if (need <= *capv && *p) return PLIFE_OK;
    dev_free(*p);
    *capv = 0;
    int rc = dev_alloc(p, need);
    if (rc) return rc;
    *capv = need;
    return PLIFE_OK;
}

int ensure_debug(plife_physics *h)
{
    if (!h->debug || h->dbg_cap >= std::max(h->n, 1)) return PLIFE_OK;
    dev_free(h->dbg_count);
    dev_free(h->dbg_hash);
    h->dbg_cap = 0;
    int rc;
    const int cap = std::max(h->n, 1);
    if ((rc = dev_alloc(&h->dbg_count, cap))) return rc;
    if ((rc = dev_alloc(&h->dbg_hash, cap))) return rc;
    h->dbg_cap = cap;
    return PLIFE_OK;
}

struct RadixPlan {
    int bits, passes, D, ntiles;
    size_t hist_len;
};

RadixPlan radix_plan(long long ncells, int n)
{
    RadixPlan p;
    p.bits = 0;
    while ((1LL << p.bits) < ncells) ++p.bits;
    if (p.bits == 0) p.bits = 1;
    p.passes = (p.bits + 7) / 8;
    p.D = (p.bits + p.passes - 1) / p.passes;
    p.ntiles = ceil_div(std::max(n, 1), kRadixTile);
    p.hist_len = (size_t)(1 << p.D) * (size_t)p.ntiles;
    return p;
}

ScanJob no_job() { return ScanJob{nullptr, 0, nullptr, nullptr, 0, nullptr}; }

// One or two device-wide scans sharing three launches.
int scan_jobs(plife_physics *h, ScanJob a, ScanJob b)
{
    if (a.n <= 0) std::swap(a, b);
    if (a.n <= 0) return PLIFE_OK;
    const int nb0 = ceil_div(a.n, kScanTile), nb1 = b.n > 0 ? ceil_div(b.n, kScanTile) : 0;
    int rc = ensure_u32(&h->bsum, &h->bsum_cap, (size_t)(nb0 + nb1));
    if (rc) return rc;
    a.bsum = h->bsum;
    b.bsum = h->bsum + nb0;
    k_scan2_reduce<<<nb0 + nb1, kBlock, 0, h->stream>>>(a, b, nb0);
    k_scan2_single<<<nb1 ? 2 : 1, kBlock, 0, h->stream>>>(a, b);
    k_scan2_down<<<nb0 + nb1, kBlock, 0, h->stream>>>(a, b, nb0);
    LAUNCH_CHECK();
    return PLIFE_OK;
}

template <int D>
int radix_pass(plife_physics *h, const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
               uint32_t *vout, int n, int shift, int ntiles, bool hist_ready)
{
    int rc;
    if (!hist_ready) {
        k_radix_hist<D><<<ntiles, kBlock, 0, h->stream>>>(kin, n, shift, ntiles, h->hist);
        if ((rc = scan_jobs(h, ScanJob{h->hist, (1 << D) * ntiles, h->hist, nullptr, 0, nullptr},
                            no_job())))
            return rc;
    }
    k_radix_scatter<D><<<ntiles, kScatterBlock, 0, h->stream>>>(kin, vin, kout, vout, h->hist, n,
                                                                 shift, ntiles);
    LAUNCH_CHECK();
    return PLIFE_OK;
}

// Stable sort of (keys0[i], i) by key -- the reference's counting-sort
// permutation (Physics.java:342-353).  The first pass's tile histogram is
// already scanned in h->hist (k_bin_tile).  Leaves the permutation (sorted
// slot -> pre-sort index) in h->perm and the sorted keys in h->skeys.
int radix_sort(plife_physics *h, int n, const RadixPlan &plan)
{
    const uint32_t *kin = h->keys0;
    const uint32_t *vin = nullptr;
    int rc;
    for (int p = 0; p < plan.passes; ++p) {
        const bool last = p == plan.passes - 1;
        uint32_t *kout = last ? h->skeys : ((p & 1) ? h->kB : h->kA);
        uint32_t *vout = (p & 1) ? h->vB : h->vA;
        const int shift = p * plan.D;
        const bool hr = p == 0;
        switch (plan.D) {
            case 1: rc = radix_pass<1>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            case 2: rc = radix_pass<2>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            case 3: rc = radix_pass<3>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            case 4: rc = radix_pass<4>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            case 5: rc = radix_pass<5>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            case 6: rc = radix_pass<6>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            case 7: rc = radix_pass<7>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
            default: rc = radix_pass<8>(h, kin, vin, kout, vout, n, shift, plan.ntiles, hr); break;
        }
        if (rc) return rc;
        kin = kout;
        vin = vout;
    }
    h->perm = const_cast<uint32_t *>(vin);
    return PLIFE_OK;
}

template <int DIM>
void launch_bin(plife_physics *h, int D, const PosBuf &P, int n, double cs, int nx, int ntiles)
{
#define PLIFE_BIN(DD)                                                                         \
    k_bin_tile<DIM, DD><<<ntiles, 1024, 0, h->stream>>>(P.x, P.y, P.z, n, cs, nx, ntiles,     \
                                                          h->keys0, h->cell_count, h->hist,   \
                                                          h->d_pairs)
    switch (D) {
        case 1: PLIFE_BIN(1); break;
        case 2: PLIFE_BIN(2); break;
        case 3: PLIFE_BIN(3); break;
        case 4: PLIFE_BIN(4); break;
        case 5: PLIFE_BIN(5); break;
        case 6: PLIFE_BIN(6); break;
        case 7: PLIFE_BIN(7); break;
        default: PLIFE_BIN(8); break;
    }
#undef PLIFE_BIN
}

struct Grid {
    int nx;
    long long ncells;
};

Grid grid_of(const plife_settings &s)
{
    Grid g;
    g.nx = (int)std::floor(1.0 / s.rmax);     // Physics.java:82-85, 312-313
    g.ncells = s.dim == 3 ? (long long)g.nx * g.nx * g.nx : (long long)g.nx * g.nx;
    return g;
}

template <int DIM, bool WRAP, bool DEBUG, bool MLDS>
void launch_force_bs(plife_physics *h, const ForceArgs &a, int bs, size_t lds)
{
    const long long grid = ((long long)a.n + bs - 1) / bs;
    switch (bs) {
        case 64: k_force<DIM, WRAP, DEBUG, MLDS, 64><<<(int)grid, 64, lds, h->stream>>>(a); break;
        case 128: k_force<DIM, WRAP, DEBUG, MLDS, 128><<<(int)grid, 128, lds, h->stream>>>(a); break;
        default: k_force<DIM, WRAP, DEBUG, MLDS, 256><<<(int)grid, 256, lds, h->stream>>>(a); break;
    }
}

template <int DIM, bool WRAP, bool DEBUG>
void launch_force(plife_physics *h, const ForceArgs &a, int bs, size_t lds)
{
    if (a.m_in_lds) launch_force_bs<DIM, WRAP, DEBUG, true>(h, a, bs, lds);
    else launch_force_bs<DIM, WRAP, DEBUG, false>(h, a, bs, lds);
}

int launch_force_any(plife_physics *h, const ForceArgs &a, int bs, size_t lds)
{
    const int dim = h->st.dim;
    const bool wrap = h->st.wrap != 0, dbg = h->debug != 0;
    if (dim == 2) {
        if (wrap) dbg ? launch_force<2, true, true>(h, a, bs, lds) : launch_force<2, true, false>(h, a, bs, lds);
        else dbg ? launch_force<2, false, true>(h, a, bs, lds) : launch_force<2, false, false>(h, a, bs, lds);
    } else {
        if (wrap) dbg ? launch_force<3, true, true>(h, a, bs, lds) : launch_force<3, true, false>(h, a, bs, lds);
        else dbg ? launch_force<3, false, true>(h, a, bs, lds) : launch_force<3, false, false>(h, a, bs, lds);
    }
    LAUNCH_CHECK();
    return PLIFE_OK;
}

// Events bracketing the phases of one step (optional).
struct PhaseEvents {
    hipEvent_t e[5];
};

int record(plife_physics *h, PhaseEvents *ev, int k)
{
    if (ev) HIP_TRY(hipEventRecord(ev->e[k], h->stream));
    return PLIFE_OK;
}

// Launch shape of the force kernel for the current world.  Explicit tuning
// fields win.  Otherwise the planner reads the particle-weighted cell
// occupancy (sum over cells of count^2, divided by N: the mean size of the
// cell a particle sits in) measured by an earlier step:
//   * uniform worlds: 256 homes per workgroup and a staging buffer that holds
//     the whole neighbourhood (one window), list capacity rising with density;
//   * clustered worlds (occupancy far above the mean N / cells): 128 homes per
//     workgroup and a small staging buffer, so that many workgroups share a CU
//     while the dense neighbourhoods stream through it in windows.
// The plan never changes results: every home visits its candidates in the
// same order whatever the windows and workgroup size.
ForcePlan plan_force(const plife_physics *h, int dim, long long ncells, int n)
{
    const int NR = dim == 3 ? 9 : 3;
    const double avg = (double)n / (double)std::max(ncells, 1LL);
    const double occ = h->occupancy > 0.0 ? h->occupancy : avg + 1.0;   // Poisson: avg + 1
    const bool clustered = h->clustered;
    const plife_tuning &t = h->tune;
    ForcePlan p;
    p.bs = t.block_homes ? t.block_homes : (clustered ? 128 : 256);
    p.lcap = t.list_cap > 0 ? std::max(8, std::min(t.list_cap, 512))
                            : (clustered ? 32 : occ >= 12.0 ? 24 : 16);
    p.lcap = std::min(p.lcap, 16384 / p.bs);             // lists stay within 32 KB of LDS
    // a workgroup of bs homes spans ~bs/avg + 1 cells of a row; it stages NR
    // rows of (span + 2) cells.  Clustered worlds keep a short cell table: the
    // few sparse workgroups that span more cells take the global path.
    const double span = p.bs / std::max(avg, 1e-9) + 1.0;
    p.max_cells = (int)std::min<double>(std::max(64.0, std::ceil(NR * (span + 2.0) * 1.5 + 16)),
                                        clustered ? 384.0 : 2048.0);
    long long want;
    size_t budget = kMaxLds;
    if (t.stage_slack > 0.0 || !clustered) {
        const double slack = t.stage_slack > 0.0 ? t.stage_slack : 1.25;
        want = (long long)std::ceil(NR * (span + 2.0) * avg * slack) + 64;
    } else {
        want = 65535;
        budget = kClusteredLds;
    }
    want = std::min<long long>((want + 63) / 64 * 64, 65535);
    const bool dbg = h->debug != 0, m_in_lds = h->ntypes <= kMaxTypesLds;
    auto total = [&](long long c) {
        return force_lds_layout(dim, dbg, h->ntypes, m_in_lds, (int)c, p.lcap, p.bs, p.max_cells).total;
    };
    // keep room for a useful staging window: shrink the cell table first
    const long long min_cap = std::min<long long>(want, 256);
    while (p.max_cells > 64 && total(min_cap) > budget) p.max_cells = std::max(64, p.max_cells * 3 / 4);
    // then shrink the staging capacity until the layout fits the LDS budget
    while (want > 64 && total(want) > budget) want -= 64;
    p.cap = (int)std::max(0LL, want);
    p.lds = force_lds_layout(dim, dbg, h->ntypes, m_in_lds, p.cap, p.lcap, p.bs, p.max_cells).total;
    return p;
}

// Read the clustering measure of the latest step the host has seen complete
// (a pinned copy refreshed by every step; a stale value only affects speed).
void refresh_occupancy(plife_physics *h)
{
    const uint64_t q = __atomic_load_n(h->stat_host, __ATOMIC_RELAXED);
    if (q == 0 || h->n <= 0) return;
    h->occupancy = (double)q / (double)h->n;
    const double avg = (double)h->n / (double)std::max(grid_of(h->st).ncells, 1LL);
    // clustered: cells seen by a particle hold far more than the average (a
    // Poisson world has occupancy avg + 1) and enough to fill long windows
    const double ratio = h->occupancy / (avg + 1.0);
    if (!h->clustered && ratio > 4.0 && h->occupancy > 64.0) h->clustered = true;
    else if (h->clustered && (ratio < 3.0 || h->occupancy < 48.0)) h->clustered = false;
}

int do_step(plife_physics *h, PhaseEvents *ev)
{
    const plife_settings &s = h->st;
    const Grid gr = grid_of(s);
    const int n = h->n;
    int rc;
    if ((rc = ensure_cells(h, gr.ncells))) return rc;
    if ((rc = ensure_debug(h))) return rc;
    const int nc = (int)gr.ncells;

    if ((rc = record(h, ev, 0))) return rc;
    const RadixPlan plan = radix_plan(gr.ncells, n);
    if ((rc = ensure_u32(&h->hist, &h->hist_cap, plan.hist_len))) return rc;
    // A step that failed between binning and scanning may leave counts behind.
    if (h->cells_dirty) {
        k_fill_u32<<<ceil_div(nc, kBlock), kBlock, 0, h->stream>>>(h->cell_count, 0u, nc);
        LAUNCH_CHECK();
    }
    h->cells_dirty = true;
    // 1. bin (Physics.java:319-340): cell keys, cell histogram, first radix
    //    digit histogram; then the cell scan (start / end offsets, counters
    //    cleared for the next step) and the digit scan in one launch
    const PosBuf P0 = h->P(h->pc);
    if (n > 0) {
        if (s.dim == 3) launch_bin<3>(h, plan.D, P0, n, s.rmax, gr.nx, plan.ntiles);
        else launch_bin<2>(h, plan.D, P0, n, s.rmax, gr.nx, plan.ntiles);
    } else {
        k_zero_u64<<<1, 64, 0, h->stream>>>(h->d_pairs, 2);
    }
    LAUNCH_CHECK();
    if ((rc = scan_jobs(h, ScanJob{h->cell_count, nc, h->cell_start, h->cell_end, 1, nullptr,
                                   h->d_pairs + 1},
                        n > 0 ? ScanJob{h->hist, (int)plan.hist_len, h->hist, nullptr, 0, nullptr}
                              : no_job())))
        return rc;
    // the clustering measure travels to pinned host memory for the launch planner
    HIP_TRY(hipMemcpyAsync(h->stat_host, h->d_pairs + 1, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           h->stream));
    h->cells_dirty = false;
    if ((rc = record(h, ev, 1))) return rc;

    h->stepped = true;
    h->last_nx = gr.nx;
    h->last_ncells = gr.ncells;
    if (n == 0) {
        h->perm = h->vA;
        for (int k = 2; k < 5; ++k)
            if ((rc = record(h, ev, k))) return rc;
        return PLIFE_OK;
    }

    // 2. stable sort of (key, index) = the reference's counting-sort permutation
    if ((rc = radix_sort(h, n, plan))) return rc;
    if ((rc = record(h, ev, 2))) return rc;
    if ((rc = record(h, ev, 3))) return rc;

    // (Physics.java:122-131 fused: the update is Jacobi)
    const double R = s.rmax * kFix;
    const double R2 = R * R;
    // 3. force + integrate: reads the pre-sort state through the permutation
    //    and writes the new state in cell order (the array swap, Physics.java:350-353)
    const int pi = h->pc, ai = h->ac, po = 1 - h->pc, ao = 1 - h->ac;
    ForceArgs a{};
    a.perm = h->perm;
    a.sx = h->px[pi];
    a.sy = h->py[pi];
    a.sz = h->pz[pi];
    a.vx = h->vx[ai];
    a.vy = h->vy[ai];
    a.vz = h->vz[ai];
    a.type = h->type[ai];
    a.id = h->id[ai];
    a.ox = h->px[po];
    a.oy = h->py[po];
    a.oz = h->pz[po];
    a.ovx = h->vx[ao];
    a.ovy = h->vy[ao];
    a.ovz = h->vz[ao];
    a.otype = h->type[ao];
    a.oid = h->id[ao];
    a.cell_start = h->cell_start;
    a.cell_end = h->cell_end;
    a.matrix = h->d_matrix;
    a.ntypes = h->ntypes;
    a.n = n;
    a.nx = gr.nx;
    const ForcePlan fp = plan_force(h, s.dim, gr.ncells, n);
    const bool m_in_lds = h->ntypes <= kMaxTypesLds;
    a.m_in_lds = m_in_lds ? 1 : 0;
    a.skeys = h->skeys;
    a.lcap = fp.lcap;
    a.max_cells = fp.max_cells;
    a.cap = fp.cap;
    const int bs = fp.bs;
    a.all_global = gr.nx < 3 ? 1 : 0;
    a.r2hi = (float)(R2 * (1.0 + 0x1p-20));
    a.r2u = (s.rmax * s.rmax) * (kFix * kFix);          // Physics.java:432, exact 2^64 scaling
    a.S = 1.0 / R;
    a.R = R;
    a.beta = s.beta;
    a.one_plus_beta = 1.0 + s.beta;
    a.inv_1mb = 1.0 / (1.0 - s.beta);
    a.inv_beta = 1.0 / s.beta;
    a.ff = std::pow(s.friction, 60 * s.dt);             // Physics.java:401
    a.kdt = s.force * s.dt / kFix;
    a.cs = s.rmax;
    a.dt = s.dt;
    a.pair_counter = h->d_pairs;
    a.dbg_count = h->dbg_count;
    a.dbg_hash = h->dbg_hash;
    const size_t lds = fp.lds;
    h->last_lds = lds;
    h->last_cap = a.cap;
    h->last_bs = bs;
    h->last_lcap = a.lcap;
    if ((rc = launch_force_any(h, a, bs, lds))) return rc;
    if ((rc = record(h, ev, 4))) return rc;
    h->pc = po;
    h->ac = ao;
    return PLIFE_OK;
}

// Allocate everything a step of the current configuration needs, so that a
// captured step allocates nothing.
int prepare_step(plife_physics *h)
{
    const Grid gr = grid_of(h->st);
    int rc;
    if ((rc = ensure_cells(h, gr.ncells))) return rc;
    if ((rc = ensure_debug(h))) return rc;
    const RadixPlan plan = radix_plan(gr.ncells, h->n);
    if ((rc = ensure_u32(&h->hist, &h->hist_cap, plan.hist_len))) return rc;
    const size_t nb = (size_t)ceil_div(gr.ncells, kScanTile) +
                      (size_t)ceil_div((long long)plan.hist_len, kScanTile);
    return ensure_u32(&h->bsum, &h->bsum_cap, nb);
}

std::vector<unsigned char> graph_key(const plife_physics *h)
{
    struct Key {
        int n, ntypes, debug, pc, ac;
        plife_settings st;
        plife_tuning tune;
        ForcePlan plan;
        const void *ptr[32];
    } k;
    std::memset(&k, 0, sizeof k);
    k.n = h->n;
    k.ntypes = h->ntypes;
    k.debug = h->debug;
    k.pc = h->pc;
    k.ac = h->ac;
    k.st = h->st;
    k.tune = h->tune;
    k.plan = plan_force(h, h->st.dim, grid_of(h->st).ncells, h->n);
    const void *p[] = {h->px[0], h->px[1], h->py[0], h->py[1], h->pz[0], h->pz[1], h->vx[0],
                       h->vx[1], h->vy[0], h->vy[1], h->vz[0], h->vz[1], h->type[0], h->type[1],
                       h->id[0], h->id[1], h->keys0, h->kA, h->kB, h->vA, h->vB, h->cell_count,
                       h->cell_start, h->cell_end, h->hist, h->bsum, h->dbg_count, h->dbg_hash,
                       h->d_matrix, h->d_pairs};
    std::memcpy(k.ptr, p, sizeof p);
    std::vector<unsigned char> v(sizeof k);
    std::memcpy(v.data(), &k, sizeof k);
    return v;
}

void drop_graph(plife_physics *h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->gkey.clear();
}

// Captures two consecutive steps into h->gexec (not executed yet).
int capture_two_steps(plife_physics *h)
{
    drop_graph(h);
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = do_step(h, nullptr);
    if (!rc) rc = do_step(h, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIP_TRY(e);
    const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        h->gexec = nullptr;
        return fail(PLIFE_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    return PLIFE_OK;
}

// nsteps steps: pairs of steps replay a captured graph, the rest launch directly.
int run_steps(plife_physics *h, int nsteps, bool honour_stop)
{
    int k = 0, rc;
    refresh_occupancy(h);
    if (nsteps >= 2 && !h->tune.direct_launch) {
        if ((rc = prepare_step(h))) return rc;
        std::vector<unsigned char> key = graph_key(h);
        if (!h->gexec || key != h->gkey) {
            if ((rc = capture_two_steps(h))) return rc;
            h->gkey = key;
        }
        while (nsteps - k >= 2) {
            if (honour_stop && h->stop.load()) return PLIFE_STOPPED;
            HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
            k += 2;
        }
    }
    for (; k < nsteps; ++k) {
        if (honour_stop && h->stop.load()) return PLIFE_STOPPED;
        if ((rc = do_step(h, nullptr))) return rc;
    }
    return PLIFE_OK;
}

int check_ready(plife_physics *h)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!h->have_settings) return fail(PLIFE_ERR_STATE, "plife_set_settings has not been called");
    return PLIFE_OK;
}

uint32_t to_fixed(double x, bool wrap)
{
    // Physics.ensurePosition (Physics.java:499-505), then round to 2^-32.
    if (!(x == x)) return 0u;
    if (wrap) {
        x = x - std::floor(x);
        double s = std::rint(x * kFix);
        return (uint32_t)(unsigned long long)s;
    }
    x = x < 0 ? 0 : (x > 1 ? 1 : x);
    double s = std::rint(x * kFix);
    return s >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)(unsigned long long)s;
}

}  // namespace

extern "C" {

const char *plife_last_error(void) { return g_last_error.c_str(); }

int plife_create(int32_t device, plife_physics **out)
{
    if (!out) return fail(PLIFE_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(PLIFE_ERR_ARGUMENT, "device %d out of range (%d devices)", device, count);
    plife_physics *h = new plife_physics();
    h->device = device;
    int rc = set_device(h);
    if (!rc) {
        hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        if (e != hipSuccess) rc = fail(PLIFE_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    if (!rc) rc = dev_alloc(&h->d_pairs, 2);
    if (!rc) {
        hipError_t e = hipHostMalloc((void **)&h->stat_host, sizeof(uint64_t), hipHostMallocDefault);
        if (e != hipSuccess) rc = fail(PLIFE_ERR_MEMORY, "hipHostMalloc: %s", hipGetErrorString(e));
        else *h->stat_host = 0;
    }
    if (!rc) rc = dev_alloc(&h->d_matrix, kMaxTypes * kMaxTypes);
    if (rc) {
        plife_destroy(h);
        return rc;
    }
    *out = h;
    g_last_error.clear();
    return PLIFE_OK;
}

int plife_destroy(plife_physics *h)
{
    if (!h) return PLIFE_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_state(h);
    dev_free(h->cell_count);
    dev_free(h->cell_start);
    dev_free(h->cell_end);
    dev_free(h->hist);
    dev_free(h->bsum);
    dev_free(h->d_pairs);
    if (h->stat_host) (void)hipHostFree(h->stat_host);
    dev_free(h->d_matrix);
    dev_free(h->dbg_count);
    dev_free(h->dbg_hash);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PLIFE_OK;
}

int plife_set_settings(plife_physics *h, const plife_settings *s, int32_t ntypes,
                       const double *matrix)
{
    if (!h || !s || !matrix) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    if (s->dim != 2 && s->dim != 3) return fail(PLIFE_ERR_ARGUMENT, "dim must be 2 or 3, got %d", s->dim);
    if (!(s->rmax > 0.0) || !(s->rmax <= 1.0))
        return fail(PLIFE_ERR_ARGUMENT, "rmax must be in (0, 1], got %g", s->rmax);
    if (!(s->beta > 0.0) || !(s->beta < 1.0))
        return fail(PLIFE_ERR_ARGUMENT, "beta must be in (0, 1), got %g", s->beta);
    if (!std::isfinite(s->dt) || !std::isfinite(s->force) || !std::isfinite(s->friction))
        return fail(PLIFE_ERR_ARGUMENT, "dt, force and friction must be finite");
    if (ntypes < 1 || ntypes > kMaxTypes)
        return fail(PLIFE_ERR_ARGUMENT, "matrix size must be in [1, %d], got %d", kMaxTypes, ntypes);
    const Grid g = grid_of(*s);
    if (g.ncells > kMaxCells)
        return fail(PLIFE_ERR_ARGUMENT, "rmax %g gives %lld cells (limit %lld)", s->rmax, g.ncells, kMaxCells);
    int rc = set_device(h);
    if (rc) return rc;
    const size_t mm = (size_t)ntypes * ntypes;
    const bool changed = ntypes != h->ntypes || h->matrix_host.size() != mm ||
                         std::memcmp(h->matrix_host.data(), matrix, mm * sizeof(double)) != 0;
    if (changed) {
        h->matrix_host.assign(matrix, matrix + mm);
        HIP_TRY(hipMemcpyAsync(h->d_matrix, h->matrix_host.data(), mm * sizeof(double),
                               hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    h->ntypes = ntypes;
    h->st = *s;
    h->have_settings = true;
    return PLIFE_OK;
}

int plife_set_particles(plife_physics *h, int32_t n, const double *pos, const double *vel,
                        const int32_t *types)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (n < 0) return fail(PLIFE_ERR_ARGUMENT, "n must be >= 0, got %d", n);
    if (n > 0 && (!pos || !vel || !types)) return fail(PLIFE_ERR_ARGUMENT, "null array");
    for (int i = 0; i < n; ++i)
        if (types[i] < 0 || types[i] >= h->ntypes)
            return fail(PLIFE_ERR_ARGUMENT, "particle %d has type %d outside [0, %d)", i, types[i], h->ntypes);
    if ((rc = set_device(h))) return rc;
    if ((rc = ensure_capacity(h, n))) return rc;
    const bool wrap = h->st.wrap != 0;
    const bool d3 = h->st.dim == 3;
    std::vector<uint32_t> x(n), y(n), z(n);
    std::vector<float> fx(n), fy(n), fz(n);
    std::vector<int32_t> ids(n);
    for (int i = 0; i < n; ++i) {
        x[i] = to_fixed(pos[3 * i], wrap);
        y[i] = to_fixed(pos[3 * i + 1], wrap);
        z[i] = d3 ? to_fixed(pos[3 * i + 2], wrap) : 0u;
        fx[i] = (float)vel[3 * i];
        fy[i] = (float)vel[3 * i + 1];
        fz[i] = d3 ? (float)vel[3 * i + 2] : 0.0f;
        ids[i] = i;
    }
    h->pc = 0;
    h->ac = 0;
    h->n = n;
    h->stepped = false;
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(h->px[0], x.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->py[0], y.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->pz[0], z.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->vx[0], fx.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->vy[0], fy.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->vz[0], fz.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->type[0], types, n * 4, hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->id[0], ids.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    // a new world: forget the clustering measure of the old one
    *h->stat_host = 0;
    h->occupancy = 0.0;
    h->clustered = false;
    return PLIFE_OK;
}

int plife_get_particles(plife_physics *h, int32_t n, double *pos, double *vel, int32_t *types,
                        int32_t *ids)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (n != h->n) return fail(PLIFE_ERR_ARGUMENT, "n = %d but the handle holds %d particles", n, h->n);
    int rc = set_device(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (n == 0) return PLIFE_OK;
    const PosBuf P = h->P(h->pc);
    const AuxBuf A = h->A(h->ac);
    const bool d3 = h->have_settings && h->st.dim == 3;
    if (pos) {
        std::vector<uint32_t> x(n), y(n), z(n);
        HIP_TRY(hipMemcpy(x.data(), P.x, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(y.data(), P.y, n * 4, hipMemcpyDeviceToHost));
        if (d3) HIP_TRY(hipMemcpy(z.data(), P.z, n * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            pos[3 * i] = (double)x[i] * kInvFix;
            pos[3 * i + 1] = (double)y[i] * kInvFix;
            pos[3 * i + 2] = d3 ? (double)z[i] * kInvFix : 0.0;
        }
    }
    if (vel) {
        std::vector<float> x(n), y(n), z(n);
        HIP_TRY(hipMemcpy(x.data(), A.vx, n * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(y.data(), A.vy, n * 4, hipMemcpyDeviceToHost));
        if (d3) HIP_TRY(hipMemcpy(z.data(), A.vz, n * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            vel[3 * i] = x[i];
            vel[3 * i + 1] = y[i];
            vel[3 * i + 2] = d3 ? (double)z[i] : 0.0;
        }
    }
    if (types) HIP_TRY(hipMemcpy(types, A.type, n * 4, hipMemcpyDeviceToHost));
    if (ids) HIP_TRY(hipMemcpy(ids, A.id, n * 4, hipMemcpyDeviceToHost));
    return PLIFE_OK;
}

int plife_get_count(plife_physics *h, int32_t *n)
{
    if (!h || !n) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    *n = h->n;
    return PLIFE_OK;
}

int plife_update(plife_physics *h, int32_t nsteps)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (nsteps < 0) return fail(PLIFE_ERR_ARGUMENT, "nsteps must be >= 0");
    if ((rc = set_device(h))) return rc;
    h->stop.store(0);
    return run_steps(h, nsteps, true);
}

int plife_force_update_stop(plife_physics *h)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    h->stop.store(1);
    return PLIFE_OK;
}

int plife_synchronize(plife_physics *h)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    int rc = set_device(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PLIFE_OK;
}

int plife_get_grid(plife_physics *h, int32_t *nx, int32_t *ncells)
{
    int rc = check_ready(h);
    if (rc) return rc;
    const Grid g = grid_of(h->st);
    if (nx) *nx = g.nx;
    if (ncells) *ncells = (int32_t)g.ncells;
    return PLIFE_OK;
}

int plife_get_cells(plife_physics *h, int32_t *keys, int32_t *cell_end, int32_t *perm,
                    int32_t ncells)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!h->stepped) return fail(PLIFE_ERR_STATE, "no step has run since the state was set");
    if (ncells != h->last_ncells)
        return fail(PLIFE_ERR_ARGUMENT, "ncells = %d but the last step used %lld cells", ncells, h->last_ncells);
    int rc = set_device(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (keys && h->n) HIP_TRY(hipMemcpy(keys, h->keys0, (size_t)h->n * 4, hipMemcpyDeviceToHost));
    if (cell_end) HIP_TRY(hipMemcpy(cell_end, h->cell_end, (size_t)ncells * 4, hipMemcpyDeviceToHost));
    if (perm && h->n) HIP_TRY(hipMemcpy(perm, h->perm, (size_t)h->n * 4, hipMemcpyDeviceToHost));
    return PLIFE_OK;
}

int plife_get_pair_count(plife_physics *h, int64_t *pairs)
{
    if (!h || !pairs) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    int rc = set_device(h);
    if (rc) return rc;
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, h->d_pairs, 8, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *pairs = (int64_t)v;
    return PLIFE_OK;
}

int plife_set_tuning(plife_physics *h, const plife_tuning *t)
{
    if (!h || !t) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    if (t->list_cap < 0 || t->list_cap > 512 || !(t->stage_slack >= 0.0) ||
        t->stage_slack > 16.0 || t->direct_launch < 0 || t->direct_launch > 1 ||
        (t->block_homes != 0 && t->block_homes != 64 && t->block_homes != 128 &&
         t->block_homes != 256))
        return fail(PLIFE_ERR_ARGUMENT, "tuning out of range");
    h->tune = *t;
    return PLIFE_OK;
}

int plife_get_tuning(plife_physics *h, plife_tuning *t, int32_t *lds_bytes)
{
    if (!h || !t) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    *t = h->tune;
    // automatic fields report what the last force launch used
    if (t->list_cap == 0) t->list_cap = h->last_lcap;
    if (t->block_homes == 0) t->block_homes = h->last_bs;
    if (lds_bytes) *lds_bytes = (int32_t)h->last_lds;
    return PLIFE_OK;
}

int plife_set_debug(plife_physics *h, int32_t on)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    h->debug = on ? 1 : 0;
    return PLIFE_OK;
}

int plife_get_neighbor_summary(plife_physics *h, int32_t n, int32_t *count, uint64_t *hash)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!h->debug || !h->stepped) return fail(PLIFE_ERR_STATE, "debug mode was off or no step has run");
    if (n != h->n) return fail(PLIFE_ERR_ARGUMENT, "n = %d but the handle holds %d particles", n, h->n);
    int rc = set_device(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (n == 0) return PLIFE_OK;
    if (count) HIP_TRY(hipMemcpy(count, h->dbg_count, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (hash) HIP_TRY(hipMemcpy(hash, h->dbg_hash, (size_t)n * 8, hipMemcpyDeviceToHost));
    return PLIFE_OK;
}

int plife_time_steps(plife_physics *h, int32_t nsteps, double *ms)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!ms || nsteps < 0) return fail(PLIFE_ERR_ARGUMENT, "bad argument");
    if ((rc = set_device(h))) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // build (or reuse) the step graph outside the timed region
    if (nsteps >= 2 && !h->tune.direct_launch) {
        if ((rc = prepare_step(h))) return rc;
        if (!h->gexec || graph_key(h) != h->gkey) {
            if ((rc = run_steps(h, 2, false))) return rc;
        }
    }
    HIP_TRY(hipEventRecord(e0, h->stream));
    rc = run_steps(h, nsteps, false);
    if (!rc) {
        HIP_TRY(hipEventRecord(e1, h->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float f = 0;
        HIP_TRY(hipEventElapsedTime(&f, e0, e1));
        *ms = f;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int plife_profile(plife_physics *h, int32_t nsteps, plife_timing *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!out || nsteps <= 0) return fail(PLIFE_ERR_ARGUMENT, "bad argument");
    if ((rc = set_device(h))) return rc;
    std::vector<PhaseEvents> ev(nsteps);
    for (auto &pe : ev)
        for (auto &e : pe.e) HIP_TRY(hipEventCreate(&e));
    for (int k = 0; k < nsteps && !rc; ++k) rc = do_step(h, &ev[k]);
    if (!rc) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        double sum[4] = {0, 0, 0, 0};
        for (auto &pe : ev)
            for (int p = 0; p < 4; ++p) {
                float f = 0;
                HIP_TRY(hipEventElapsedTime(&f, pe.e[p], pe.e[p + 1]));
                sum[p] += f;
            }
        float wall = 0;
        HIP_TRY(hipEventElapsedTime(&wall, ev.front().e[0], ev.back().e[4]));
        out->bin_ms = sum[0] / nsteps;
        out->sort_ms = sum[1] / nsteps;
        out->gather_ms = sum[2] / nsteps;
        out->force_ms = sum[3] / nsteps;
        out->step_ms = wall / nsteps;
        out->steps = nsteps;
        int64_t pairs = 0;
        rc = plife_get_pair_count(h, &pairs);
        out->pairs = pairs;
    }
    for (auto &pe : ev)
        for (auto &e : pe.e) (void)hipEventDestroy(e);
    return rc;
}

}  // extern "C"
