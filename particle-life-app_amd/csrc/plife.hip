// plife.hip -- C ABI (include/plife.h) of the gfx950 particle-life physics step.
//
// Host orchestration of Physics.update() (backend/Physics.java:112-134) for a
// batch of independent worlds; the single-world handle is a batch of one:
//   bin (cell keys + cell histogram + first digit histogram) -> scans ->
//   stable radix sort -> force + integrate (reads the pre-sort state through
//   the permutation, writes the cell-sorted new state),
// all enqueued on the handle's own stream.  No torch, no CPU fallback: every
// step runs the HIP kernels of plife_kernels.hpp and plife_force.hpp.
//
// Batch layout: world w owns particle slots [off_w, off_w + n_w) and cells
// [cell_base_w, cell_base_w + nx_w^dim).  Keys of different worlds never
// interleave, so one stable sort of the batch keys leaves every world in its
// own slot range, sorted exactly as stepping it alone sorts it; force
// workgroups never straddle two worlds and every home visits its candidates
// in the same order, so each world's results are bit-identical to stepping it
// alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sched.h>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/plife.h"
#include "plife_force.hpp"
#include "plife_small.hpp"

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
constexpr int kMaxWorlds = 65535;     // the world index of a slot is a u16
constexpr int kMaxLds = 65536;
constexpr int kClusteredLds = 17152;  // force-kernel LDS in clustered worlds: 9 workgroups per CU
constexpr int kTuneSteps = 24;         // an update long enough to time two configurations
constexpr int kTuneMax = 22;           // steps such a timing takes at most (4 x 4 + 3 switches + 3)
constexpr size_t kWordLds = 24576;    // most LDS for the balanced path's mask words

struct ForcePlan {
    int wpool = 0;                    // mask words per workgroup of the balanced path (0: windowed only)
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

// Capacity for `need` elements when an array of capacity `cap` is too small:
// exact the first time, then at least 1.5x, so that a batch growing one world
// at a time reallocates O(log n) times (hipFree waits for the whole device,
// including other handles' streams).
inline size_t grown(size_t cap, size_t need) { return cap ? std::max(need, cap + cap / 2) : need; }

// Arrays of at most this size that a larger one replaces are kept until the
// handle is destroyed instead of freed at once: hipFree waits for the whole
// device, and with 1.5x growth the kept arrays total less than twice the
// live one.
constexpr size_t kRetireMax = (size_t)32 << 20;

// One pinned staging buffer of table uploads and the event recorded after the
// copies that read it.
struct UploadSlot {
    unsigned char *p = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};
constexpr int kUploadSlots = 4;
constexpr int kDlChunks = 8;           // downloads move in at most this many chunks
constexpr int kFlightSlots = 64;       // events marking queued launches of plife_update
constexpr double kStopWindowMs = 20.0; // device work plife_update keeps queued (stop latency)
constexpr int kDlMinChunk = 65536;     // slots; smaller downloads are one chunk

// Grow-only device array: reallocated (contents dropped) when too small.  The
// old array goes to `retired` when given and small enough, else is freed.
template <typename T>
int ensure_dev(T **p, size_t *capv, size_t need, std::vector<void *> *retired = nullptr)
{
    need = std::max<size_t>(need, 1);
    if (need <= *capv && *p) return PLIFE_OK;
    const size_t cap = grown(*p ? *capv : 0, need);
    if (*p && retired && *capv * sizeof(T) <= kRetireMax) {
        retired->push_back((void *)*p);
        *p = nullptr;
    }
    dev_free(*p);
    *capv = 0;
    int rc = dev_alloc(p, cap);
    if (rc) return rc;
    *capv = cap;
    return PLIFE_OK;
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

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

}  // namespace

// One world as the host keeps it.
struct WorldHost {
    int32_t id = 0;
    plife_settings st{};
    int ntypes = 0;
    std::vector<double> matrix;
    int n = 0, off = 0;
    int64_t steps = 0;
    bool live = true;                  // false once removed (dropped at the next layout)
    double nb = -1.0;                  // neighbours per particle in the last measured step
    int max_type = -1;                 // highest type of its particles (must stay < ntypes to step)
    bool sparse = false;               // stepped by the sparse-world force kernel
};

struct plife_physics {
    int device = 0;
    hipStream_t stream = nullptr;
    bool batch = false;                // created by plife_batch_create
    bool have_settings = false;        // single-world handle: plife_set_settings was called
    int dim = 0;                       // shared by all worlds

    // worlds and their device layout
    std::vector<WorldHost> worlds;
    std::unordered_map<int32_t, int> index_of;   // world id -> index in worlds
    int32_t next_id = 0;
    int used = 0;                      // slots of the current buffers holding particles
    bool layout_dirty = true;          // composition or particle counts changed
    bool params_dirty = true;          // settings or matrices changed
    uint64_t version = 0;              // bumped whenever the uploaded layout changes
    std::vector<WorldDev> wdev;        // host copy of d_worlds
    WorldDev *d_worlds = nullptr;
    size_t worlds_cap = 0;
    uint16_t *world_of = nullptr;      // world index of each slot (batches of several worlds)
    size_t world_of_cap = 0;
    int *d_src_off = nullptr;          // compaction source offsets
    size_t src_off_cap = 0;
    std::vector<double> matrices;      // all worlds' matrices, concatenated
    double *d_matrix = nullptr;
    size_t matrix_cap = 0;
    long long ncells = 0;              // all worlds
    int ntypes_max = 0;

    // force workgroups: (world, first home) per block; periodic worlds first
    std::vector<uint2> blocks;
    uint2 *d_blocks = nullptr;
    size_t blocks_cap = 0;
    int blocks_bs = 0, nb_total = 0;
    int seg[4] = {};                   // blocks of (periodic, clamped) x (staged, sparse) worlds
    uint64_t blocks_version = ~0ull, blocks_classes = ~0ull;
    uint64_t blocks_serial = 0;        // bumped whenever the table's content or order changes
    std::vector<uint64_t> block_key;   // per entry: world id, sparse kernel, first home in the world
    std::vector<uint32_t> block_est;   // per entry: expected pairs (before a step measures them)
    uint32_t *work_host = nullptr;     // pinned: each block's pairs, written by the force kernels
    uint32_t *work_dev = nullptr;      // the device's address of work_host
    size_t work_host_cap = 0;
    uint64_t classes = 0;              // bumped whenever a world changes force kernel

    // particle state: positions ping-pong P[2], the rest A[2]
    int n = 0, cap = 0;
    uint32_t *px[2] = {}, *py[2] = {}, *pz[2] = {};
    float *vx[2] = {}, *vy[2] = {}, *vz[2] = {};
    int32_t *type[2] = {}, *id[2] = {};
    int pc = 0, ac = 0;

    // binning / sort scratch
    uint32_t *keys0 = nullptr, *kA = nullptr, *kB = nullptr, *vA = nullptr, *vB = nullptr;
    uint32_t *skeys = nullptr;         // batch cell key of each sorted slot (last radix pass)
    uint32_t *perm = nullptr;          // points at vA or vB after a step
    size_t cells_cap = 0;
    uint32_t *cell_count = nullptr, *cell_start = nullptr, *cell_end = nullptr;
    size_t hist_cap = 0, bsum_cap = 0;
    uint32_t *hist = nullptr, *bsum = nullptr;
    unsigned long long *bsq = nullptr;   // per-block sums of squares of the cell scan
    size_t bsq_cap = 0;
    // [0] interacting pairs, [1] sum of cell counts^2, [2 + w] pairs of world w
    unsigned long long *d_pairs = nullptr;
    size_t pairs_cap = 0;
    // pinned copy of d_pairs (3 + worlds entries), written by every step
    uint64_t *stat_host = nullptr;
    unsigned long long *stat_dev = nullptr;   // the device's address of stat_host
    size_t stat_cap = 0;
    std::vector<void *> retired_host;  // replaced pinned buffers, freed at destroy
    unsigned char *pin = nullptr;      // pinned staging of state downloads
    size_t pin_cap = 0;
    uint32_t *dl_dev = nullptr;        // device staging of state downloads (chunk-major)
    size_t dl_cap = 0;
    hipEvent_t dl_ev[kDlChunks] = {};  // one per download chunk: its DMA has completed
    UploadSlot up[kUploadSlots];       // pinned staging of table uploads, used in turn
    int up_cur = 0;
    size_t up_used = 0;
    std::vector<void *> retired;       // replaced small device arrays, freed at destroy
    SlotEdit *d_edits = nullptr;       // in-place particle edits of the last plife_set_particle_slots
    size_t edits_cap = 0;
    double occupancy = 0.0;            // particle-weighted cell occupancy seen by the planner
    bool clustered = false;            // the planner's regime (with hysteresis)

    // the last step
    bool stepped = false;
    long long last_ncells = 0;

    // debug
    int debug = 0;
    size_t dbg_cap = 0, dbg_hcap = 0;
    int32_t *dbg_count = nullptr;
    unsigned long long *dbg_hash = nullptr;
    float *dbg_acc = nullptr;
    size_t dbg_acap = 0;

    std::atomic<int> stop{0};
    bool cells_dirty = false;          // cell_count may hold stale counts

    // launches queued by the running plife_update: a ring of events, each
    // marking the end of one launch of fl_steps[] steps
    hipEvent_t fl_ev[kFlightSlots] = {};
    int fl_steps[kFlightSlots] = {};
    int fl_head = 0, fl_count = 0, fl_queued = 0;

    // HIP graph of two consecutive steps (buffer roles repeat every 2 steps)
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned char> gkey;
    std::vector<unsigned char> seen_key;   // configuration of the previous update

    // launch tuning (0 = automatic), the current plan, what the last launch used
    plife_tuning tune{};
    ForcePlan plan{};
    size_t last_lds = 0;
    int last_cap = 0, last_bs = 0, last_lcap = 0;

    // the one-launch step of worlds that fit one workgroup's LDS (k_step_small)
    int lds_limit = 65536;             // the device's LDS per workgroup
    bool small_ok = false;             // every world fits and the tuning allows the one-launch step
    bool small = false;                // the current configuration steps in one launch
    bool small_pref = false;           // the measured choice for layout version small_meas
    uint64_t small_meas = ~0ull;
    hipEvent_t tune_ev[2] = {};        // timing of the measured choice
    int small_homes = 256, small_nmax = 0, small_cmax = 0;
    size_t small_lds = 0;
    int small_kbuf = 0;                // which of skeys (0) and kA (1) holds the last step's sorted keys
    unsigned long long *d_facc = nullptr;   // its counter accumulator (+ ticket), zero between steps
    size_t facc_cap = 0;

    // homes per force workgroup of the multi-launch path where block_homes is
    // automatic: measured once per class of layout (shape_class) and kept
    int auto_bs = 0;                   // the choice for layout version bs_meas (0: the planner's own)
    uint64_t bs_meas = ~0ull;
    std::vector<std::pair<uint64_t, int>> bs_seen;   // (class, homes) of the classes measured

    PosBuf P(int k) { return PosBuf{px[k], py[k], pz[k]}; }
    AuxBuf A(int k) { return AuxBuf{vx[k], vy[k], vz[k], type[k], id[k]}; }
};

struct plife_batch : plife_physics {};

namespace {

int set_device(plife_physics *h)
{
    HIP_TRY(hipSetDevice(h->device));
    return PLIFE_OK;
}

void free_state_arrays(uint32_t *px[2], uint32_t *py[2], uint32_t *pz[2], float *vx[2],
                       float *vy[2], float *vz[2], int32_t *type[2], int32_t *id[2])
{
    for (int k = 0; k < 2; ++k) {
        dev_free(px[k]);
        dev_free(py[k]);
        dev_free(pz[k]);
        dev_free(vx[k]);
        dev_free(vy[k]);
        dev_free(vz[k]);
        dev_free(type[k]);
        dev_free(id[k]);
    }
}

void free_state(plife_physics *h)
{
    free_state_arrays(h->px, h->py, h->pz, h->vx, h->vy, h->vz, h->type, h->id);
    dev_free(h->keys0);
    dev_free(h->kA);
    dev_free(h->kB);
    dev_free(h->vA);
    dev_free(h->vB);
    dev_free(h->skeys);
    h->perm = nullptr;
    h->cap = 0;
}

// Room for `need` particles.  With keep, the particles in the first h->used
// slots of the current buffers move to the new buffers.
int ensure_capacity(plife_physics *h, int need, bool keep)
{
    if (need <= h->cap && h->cap > 0) return PLIFE_OK;
    const int cap = std::max({need, 1, keep ? h->cap + h->cap / 2 : 0});
    uint32_t *px[2] = {}, *py[2] = {}, *pz[2] = {};
    float *vx[2] = {}, *vy[2] = {}, *vz[2] = {};
    int32_t *type[2] = {}, *id[2] = {};
    int rc = PLIFE_OK;
    for (int k = 0; k < 2 && !rc; ++k) {
        if (!rc) rc = dev_alloc(&px[k], cap);
        if (!rc) rc = dev_alloc(&py[k], cap);
        if (!rc) rc = dev_alloc(&pz[k], cap);
        if (!rc) rc = dev_alloc(&vx[k], cap);
        if (!rc) rc = dev_alloc(&vy[k], cap);
        if (!rc) rc = dev_alloc(&vz[k], cap);
        if (!rc) rc = dev_alloc(&type[k], cap);
        if (!rc) rc = dev_alloc(&id[k], cap);
    }
    if (rc) {
        free_state_arrays(px, py, pz, vx, vy, vz, type, id);
        return rc;
    }
    hipError_t e = hipSuccess;
    if (keep && h->used > 0) {
        const size_t b = (size_t)h->used * 4;
        const int p = h->pc, a = h->ac;
        const hipMemcpyKind k = hipMemcpyDeviceToDevice;
        if (e == hipSuccess) e = hipMemcpyAsync(px[0], h->px[p], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(py[0], h->py[p], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pz[0], h->pz[p], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(vx[0], h->vx[a], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(vy[0], h->vy[a], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(vz[0], h->vz[a], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(type[0], h->type[a], b, k, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(id[0], h->id[a], b, k, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        free_state_arrays(px, py, pz, vx, vy, vz, type, id);
        return fail(PLIFE_ERR_DEVICE, "state copy failed: %s", hipGetErrorString(e));
    }
    free_state(h);
    for (int k = 0; k < 2; ++k) {
        h->px[k] = px[k];
        h->py[k] = py[k];
        h->pz[k] = pz[k];
        h->vx[k] = vx[k];
        h->vy[k] = vy[k];
        h->vz[k] = vz[k];
        h->type[k] = type[k];
        h->id[k] = id[k];
    }
    h->pc = 0;
    h->ac = 0;
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
    const size_t need = (size_t)std::max(ncells, 1LL);
    if (need <= h->cells_cap && h->cell_count) return PLIFE_OK;
    const size_t c = grown(h->cell_count ? h->cells_cap : 0, need);
    if (h->cell_count && h->cells_cap * 4 <= kRetireMax) {
        h->retired.insert(h->retired.end(), {h->cell_count, h->cell_start, h->cell_end});
        h->cell_count = h->cell_start = h->cell_end = nullptr;
    }
    dev_free(h->cell_count);
    dev_free(h->cell_start);
    dev_free(h->cell_end);
    h->cells_cap = 0;
    int rc;
    if ((rc = dev_alloc(&h->cell_count, c))) return rc;
    if ((rc = dev_alloc(&h->cell_start, c))) return rc;
    if ((rc = dev_alloc(&h->cell_end, c))) return rc;
    // the step keeps cell_count all-zero between steps (the scan clears it)
    HIP_TRY(hipMemsetAsync(h->cell_count, 0, c * 4, h->stream));
    h->cells_cap = c;
    return PLIFE_OK;
}

int ensure_debug(plife_physics *h)
{
    if (!h->debug) return PLIFE_OK;
    const size_t need = (size_t)std::max(h->n, 1);
    int rc;
    if ((rc = ensure_dev(&h->dbg_count, &h->dbg_cap, need))) return rc;
    if ((rc = ensure_dev(&h->dbg_acc, &h->dbg_acap, 3 * need))) return rc;
    return ensure_dev(&h->dbg_hash, &h->dbg_hcap, need);
}

// ---------------------------------------------------------------------------
// world layout
// ---------------------------------------------------------------------------

WorldDev world_dev(const WorldHost &w, uint32_t cell_base, int moff)
{
    const plife_settings &s = w.st;
    const Grid gr = grid_of(s);
    const double R = s.rmax * kFix;
    WorldDev d{};
    d.off = w.off;
    d.n = w.n;
    d.nx = gr.nx;
    d.all_global = gr.nx < 3 ? 1 : 0;   // repeated stencil cells (Physics.java:377-395)
    d.cell_base = cell_base;
    d.ntypes = w.ntypes;
    d.moff = moff;
    d.wrap = s.wrap ? 1 : 0;
    d.r2hi = (float)(R * R * (1.0 + 0x1p-20));
    d.cs = s.rmax;
    d.r2u = (s.rmax * s.rmax) * (kFix * kFix);         // Physics.java:432, exact 2^64 scaling
    d.S = 1.0 / R;
    d.R = R;
    d.beta = s.beta;
    d.one_plus_beta = 1.0 + s.beta;
    d.inv_1mb = 1.0 / (1.0 - s.beta);
    d.inv_beta = 1.0 / s.beta;
    d.m2S = -2.0 * d.S;
    d.betaR = s.beta * R;
    d.ff = std::pow(s.friction, 60 * s.dt);            // Physics.java:401
    d.ka = s.force * kInvFix;
    d.kdt = s.force * s.dt / kFix;
    d.dt = s.dt;
    return d;
}

// Table uploads go through pinned buffers and are ordered on the stream, so
// the host does not wait for them.  A batch of uploads (begin_uploads with
// room for `bytes`, upload calls, end_uploads) takes the next of
// kUploadSlots buffers, waiting only until the batch that used it last has
// been copied.
int begin_uploads(plife_physics *h, size_t bytes)
{
    h->up_cur = (h->up_cur + 1) % kUploadSlots;
    UploadSlot &s = h->up[h->up_cur];
    HIP_TRY(hipEventSynchronize(s.done));
    h->up_used = 0;
    if (bytes <= s.cap) return PLIFE_OK;
    const size_t cap = std::max({bytes, s.cap * 2, (size_t)64 * 1024});
    if (s.p) (void)hipHostFree(s.p);
    s.p = nullptr;
    s.cap = 0;
    const hipError_t e = hipHostMalloc((void **)&s.p, cap, hipHostMallocDefault);
    if (e != hipSuccess) {
        s.p = nullptr;
        return fail(PLIFE_ERR_MEMORY, "hipHostMalloc of %zu bytes failed: %s", cap, hipGetErrorString(e));
    }
    s.cap = cap;
    return PLIFE_OK;
}

int upload(plife_physics *h, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return PLIFE_OK;
    UploadSlot &s = h->up[h->up_cur];
    const size_t at = (h->up_used + 255) & ~(size_t)255;
    if (at + bytes > s.cap) return fail(PLIFE_ERR_STATE, "upload staging too small");
    std::memcpy(s.p + at, src, bytes);
    h->up_used = at + bytes;
    HIP_TRY(hipMemcpyAsync(dst, s.p + at, bytes, hipMemcpyHostToDevice, h->stream));
    return PLIFE_OK;
}

int end_uploads(plife_physics *h)
{
    HIP_TRY(hipEventRecord(h->up[h->up_cur].done, h->stream));
    return PLIFE_OK;
}

// Rebuilds the world table (offsets, cell bases, matrix offsets) from
// h->worlds and uploads it with the matrices, leaving `extra` bytes of the
// upload staging for the caller's own uploads in the same batch.
int upload_params(plife_physics *h, size_t extra)
{
    const int W = (int)h->worlds.size();
    h->wdev.resize(W);
    h->matrices.clear();
    long long cells = 0;
    int tmax = 0;
    for (int w = 0; w < W; ++w) {
        const WorldHost &wh = h->worlds[w];
        h->wdev[w] = world_dev(wh, (uint32_t)cells, (int)h->matrices.size());
        h->matrices.insert(h->matrices.end(), wh.matrix.begin(), wh.matrix.end());
        cells += grid_of(wh.st).ncells;
        tmax = std::max(tmax, wh.ntypes);
    }
    if (cells > kMaxCells)
        return fail(PLIFE_ERR_ARGUMENT, "the worlds need %lld cells together (limit %lld)", cells, kMaxCells);
    h->ncells = cells;
    h->ntypes_max = tmax;
    int rc;
    if ((rc = ensure_dev(&h->d_worlds, &h->worlds_cap, (size_t)W, &h->retired))) return rc;
    if ((rc = ensure_dev(&h->d_matrix, &h->matrix_cap, h->matrices.size(), &h->retired))) return rc;
    unsigned long long *const old_pairs = h->d_pairs;
    if ((rc = ensure_dev(&h->d_pairs, &h->pairs_cap, (size_t)(3 + W), &h->retired))) return rc;
    if (h->d_pairs != old_pairs)        // per-world counters read before the next step
        HIP_TRY(hipMemsetAsync(h->d_pairs, 0, h->pairs_cap * 8, h->stream));
    unsigned long long *const old_facc = h->d_facc;
    if ((rc = ensure_dev(&h->d_facc, &h->facc_cap, (size_t)(3 + W), &h->retired))) return rc;
    if (h->d_facc != old_facc)          // the one-launch step keeps it zero between steps
        HIP_TRY(hipMemsetAsync(h->d_facc, 0, h->facc_cap * 8, h->stream));
    if ((size_t)(3 + W) > h->stat_cap) {
        // the old buffer may still be written by queued steps: keep it until destroy
        const size_t cap = grown(h->stat_cap, (size_t)(3 + W));
        uint64_t *p = nullptr;
        const hipError_t e = hipHostMalloc((void **)&p, cap * 8, hipHostMallocDefault);
        if (e != hipSuccess) return fail(PLIFE_ERR_MEMORY, "hipHostMalloc: %s", hipGetErrorString(e));
        std::memset(p, 0, cap * 8);
        p[1] = h->stat_host ? h->stat_host[1] : 0;
        unsigned long long *dp = nullptr;
        const hipError_t e2 = hipHostGetDevicePointer((void **)&dp, p, 0);
        if (e2 != hipSuccess) {
            (void)hipHostFree(p);
            return fail(PLIFE_ERR_DEVICE, "hipHostGetDevicePointer: %s", hipGetErrorString(e2));
        }
        if (h->stat_host) h->retired_host.push_back(h->stat_host);
        h->stat_host = p;
        h->stat_dev = dp;
        h->stat_cap = cap;
    }
    const size_t wb = W * sizeof(WorldDev), mb = h->matrices.size() * sizeof(double);
    if ((rc = begin_uploads(h, wb + mb + extra + 1024))) return rc;
    if ((rc = upload(h, h->d_worlds, h->wdev.data(), wb))) return rc;
    return upload(h, h->d_matrix, h->matrices.data(), mb);
}

// Brings the device layout up to date: drops removed worlds (one compaction
// pass into the other state buffers), assigns offsets, uploads the world
// table and the per-slot world index.  Called outside graph capture.
int sync_layout(plife_physics *h)
{
    if (!h->layout_dirty && !h->params_dirty) return PLIFE_OK;
    int rc;
    if (h->layout_dirty) {
        // old offsets of the surviving worlds, then the new dense layout
        std::vector<WorldHost> kept;
        std::vector<int> src;
        bool moved = false, reindexed = false;
        int off = 0;
        for (WorldHost &w : h->worlds) {
            if (!w.live) {
                reindexed = true;
                continue;
            }
            src.push_back(w.off);
            if (w.off != off) moved = true;
            w.off = off;
            off += w.n;
            kept.push_back(std::move(w));
        }
        h->worlds = std::move(kept);
        h->index_of.clear();
        for (int w = 0; w < (int)h->worlds.size(); ++w) h->index_of[h->worlds[w].id] = w;
        const int W = (int)h->worlds.size();
        const int nnew = off;
        if ((rc = upload_params(h, W * sizeof(int)))) return rc;
        if (reindexed)      // per-world pair counters of the last step refer to old indices
            HIP_TRY(hipMemsetAsync(h->d_pairs, 0, (size_t)(3 + W) * 8, h->stream));
        if (W > 1 && nnew > 0) {
            if ((rc = ensure_dev(&h->world_of, &h->world_of_cap, (size_t)nnew, &h->retired))) return rc;
            k_world_of<<<ceil_div(nnew, kBlock), kBlock, 0, h->stream>>>(nnew, h->d_worlds, W,
                                                                        h->world_of);
            LAUNCH_CHECK();
        }
        if (moved && nnew > 0) {
            if ((rc = ensure_dev(&h->d_src_off, &h->src_off_cap, (size_t)W, &h->retired))) return rc;
            if ((rc = upload(h, h->d_src_off, src.data(), W * sizeof(int)))) return rc;
            const int p = h->pc, a = h->ac;
            k_compact<<<ceil_div(nnew, kBlock), kBlock, 0, h->stream>>>(
                nnew, h->d_worlds, W > 1 ? h->world_of : nullptr, h->d_src_off, h->px[p], h->py[p],
                h->pz[p], h->vx[a], h->vy[a], h->vz[a], h->type[a], h->id[a], h->px[1 - p],
                h->py[1 - p], h->pz[1 - p], h->vx[1 - a], h->vy[1 - a], h->vz[1 - a], h->type[1 - a],
                h->id[1 - a]);
            LAUNCH_CHECK();
            h->pc = 1 - p;
            h->ac = 1 - a;
        }
        if ((rc = end_uploads(h))) return rc;
        h->n = nnew;
        h->used = nnew;
        h->stepped = false;
    } else {
        if ((rc = upload_params(h, 0))) return rc;
        if ((rc = end_uploads(h))) return rc;
    }
    h->layout_dirty = false;
    h->params_dirty = false;
    ++h->version;
    return PLIFE_OK;
}

// Force workgroup table: segments of periodic staged, periodic sparse,
// clamped staged and clamped sparse worlds (one launch each), bs homes per
// staged block and kSparseBlock per sparse block.
int ensure_blocks(plife_physics *h, int bs)
{
    if (h->blocks_bs == bs && h->blocks_version == h->version && h->blocks_classes == h->classes)
        return PLIFE_OK;
    // the last step's work of the blocks that survive, by block identity
    std::unordered_map<uint64_t, uint32_t> prev;
    const int nold = (int)h->blocks.size();
    if (nold > 0 && h->work_host && h->work_host_cap > (size_t)nold &&
        (int)h->block_key.size() == nold &&
        __atomic_load_n(h->work_host + nold, __ATOMIC_RELAXED) == (uint32_t)(h->blocks_serial + 1))
        for (int i = 0; i < nold; ++i)
            prev[h->block_key[i]] = __atomic_load_n(h->work_host + i, __ATOMIC_RELAXED);
    h->blocks.clear();
    h->block_key.clear();
    h->block_est.clear();
    for (int s = 0; s < 4; ++s) {
        const bool wrap = s < 2, sparse = (s & 1) != 0;
        const size_t before = h->blocks.size();
        for (int w = 0; w < (int)h->worlds.size(); ++w) {
            const WorldHost &wh = h->worlds[w];
            if ((wh.st.wrap != 0) != wrap || wh.sparse != sparse) continue;
            const int step = sparse ? kSparseBlock : bs;
            // expected pairs per home where no measurement carries over
            const double r = wh.st.rmax;
            const double ball = wh.st.dim == 3 ? 4.0 / 3.0 * M_PI * r * r * r : M_PI * r * r;
            const double nb = wh.nb >= 0.0 ? wh.nb : std::max(wh.n - 1, 0) * ball;
            for (int g = 0; g < wh.n; g += step) {
                const uint64_t key = ((uint64_t)(uint32_t)wh.id << 33) | ((uint64_t)sparse << 32) | (uint32_t)g;
                const auto it = prev.find(key);
                const double homes = std::min(step, wh.n - g);
                h->blocks.push_back(make_uint2((uint32_t)w, (uint32_t)(wh.off + g)));
                h->block_key.push_back(key);
                h->block_est.push_back(it != prev.end() ? it->second
                                                        : (uint32_t)std::min(nb * homes, 4.0e9));
            }
        }
        h->seg[s] = (int)(h->blocks.size() - before);
    }
    h->nb_total = (int)h->blocks.size();
    int rc;
    if ((rc = ensure_dev(&h->d_blocks, &h->blocks_cap, h->blocks.size(), &h->retired))) return rc;
    if (h->blocks.size() + 1 > h->work_host_cap) {
        // the old buffer may still be written by queued steps: keep it until destroy
        const size_t cap = grown(h->work_host_cap, h->blocks.size() + 1);
        uint32_t *p = nullptr;
        const hipError_t e = hipHostMalloc((void **)&p, cap * 4, hipHostMallocDefault);
        if (e != hipSuccess) return fail(PLIFE_ERR_MEMORY, "hipHostMalloc: %s", hipGetErrorString(e));
        std::memset(p, 0, cap * 4);
        uint32_t *dp = nullptr;
        const hipError_t e2 = hipHostGetDevicePointer((void **)&dp, p, 0);
        if (e2 != hipSuccess) {
            (void)hipHostFree(p);
            return fail(PLIFE_ERR_DEVICE, "hipHostGetDevicePointer: %s", hipGetErrorString(e2));
        }
        if (h->work_host) h->retired_host.push_back(h->work_host);
        h->work_host = p;
        h->work_dev = dp;
        h->work_host_cap = cap;
    }
    ++h->blocks_serial;
    if (!h->blocks.empty()) {
        const size_t bytes = h->blocks.size() * sizeof(uint2);
        if ((rc = begin_uploads(h, bytes))) return rc;
        if ((rc = upload(h, h->d_blocks, h->blocks.data(), bytes))) return rc;
        if ((rc = end_uploads(h))) return rc;
    }
    h->blocks_bs = bs;
    h->blocks_version = h->version;
    h->blocks_classes = h->classes;
    return PLIFE_OK;
}

// ---------------------------------------------------------------------------
// scans, sort, binning, force launches
// ---------------------------------------------------------------------------

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

ScanJob no_job() { return ScanJob{nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr}; }

// One or two device-wide scans sharing three launches.
int scan_jobs(plife_physics *h, ScanJob a, ScanJob b)
{
    if (a.n <= 0) std::swap(a, b);
    if (a.n <= 0) return PLIFE_OK;
    const int nb0 = ceil_div(a.n, kScanTile), nb1 = b.n > 0 ? ceil_div(b.n, kScanTile) : 0;
    if (a.n <= kSmallScan && b.n <= kSmallScan) {
        k_scan_small<<<nb1 ? 2 : 1, kSmallBlock, 0, h->stream>>>(a, b);
        LAUNCH_CHECK();
        return PLIFE_OK;
    }
    if ((size_t)(nb0 + nb1) > h->bsum_cap || !h->bsum)
        return fail(PLIFE_ERR_STATE, "scan scratch was not prepared");
    if ((a.sq && (!h->bsq || (size_t)nb0 > h->bsq_cap)) || b.sq)
        return fail(PLIFE_ERR_STATE, "scan scratch was not prepared");
    a.bsum = h->bsum;
    b.bsum = h->bsum + nb0;
    a.sqb = h->bsq;
    k_scan2_reduce<<<nb0 + nb1, kBlock, 0, h->stream>>>(a, b, nb0);
    if (nb0 <= kFoldBlocks && nb1 <= kFoldBlocks) {
        k_scan2_down_fold<<<nb0 + nb1, kBlock, 0, h->stream>>>(a, b, nb0);
        LAUNCH_CHECK();
        return PLIFE_OK;
    }
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
        if ((rc = scan_jobs(h, ScanJob{h->hist, (1 << D) * ntiles, h->hist, nullptr, 0, nullptr, nullptr},
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
void launch_bin(plife_physics *h, int D, const PosBuf &P, int n, int ntiles)
{
    const uint16_t *wof = h->worlds.size() > 1 ? h->world_of : nullptr;
    const int ncounters = 2 + (int)h->worlds.size();
#define PLIFE_BIN(DD)                                                                          \
    k_bin_tile<DIM, DD><<<ntiles, 1024, 0, h->stream>>>(P.x, P.y, P.z, n, h->d_worlds, wof,    \
                                                          ntiles, h->keys0, h->cell_count,     \
                                                          h->hist, h->d_pairs, ncounters,      \
                                                          h->version + 1, h->stat_dev)
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

template <int DIM, bool WRAP, bool DEBUG, bool MLDS>
void launch_force_bs(plife_physics *h, const ForceArgs &a, int nblocks, int bs, size_t lds)
{
    switch (bs) {
        case 64: k_force<DIM, WRAP, DEBUG, MLDS, 64><<<nblocks, 64, lds, h->stream>>>(a); break;
        case 128: k_force<DIM, WRAP, DEBUG, MLDS, 128><<<nblocks, 128, lds, h->stream>>>(a); break;
        default: k_force<DIM, WRAP, DEBUG, MLDS, 256><<<nblocks, 256, lds, h->stream>>>(a); break;
    }
}

template <int DIM, bool WRAP, bool DEBUG>
void launch_force(plife_physics *h, const ForceArgs &a, int nblocks, int bs, size_t lds)
{
    if (a.m_in_lds) launch_force_bs<DIM, WRAP, DEBUG, true>(h, a, nblocks, bs, lds);
    else launch_force_bs<DIM, WRAP, DEBUG, false>(h, a, nblocks, bs, lds);
}

template <int DIM, bool WRAP, bool DEBUG>
void launch_sparse(plife_physics *h, const ForceArgs &a, int nblocks)
{
    const size_t lds = a.m_in_lds ? (size_t)a.ntypes_max * a.ntypes_max * 8 : 0;
    if (a.m_in_lds) k_force_sparse<DIM, WRAP, DEBUG, true><<<nblocks, kSparseBlock, lds, h->stream>>>(a);
    else k_force_sparse<DIM, WRAP, DEBUG, false><<<nblocks, kSparseBlock, 0, h->stream>>>(a);
}

// One launch per segment of the block table present in the batch: periodic
// staged, periodic sparse, clamped staged, clamped sparse.
int launch_force_any(plife_physics *h, ForceArgs a, int bs, size_t lds)
{
    const bool dbg = h->debug != 0;
    int base = 0;
    for (int s = 0; s < 4; ++s) {
        const bool wrap = s < 2, sparse = (s & 1) != 0;
        const int nb = h->seg[s];
        a.block_base = base;
        base += nb;
        if (nb <= 0) continue;
        if (sparse) {
            if (h->dim == 3) {
                if (wrap) dbg ? launch_sparse<3, true, true>(h, a, nb) : launch_sparse<3, true, false>(h, a, nb);
                else dbg ? launch_sparse<3, false, true>(h, a, nb) : launch_sparse<3, false, false>(h, a, nb);
            } else {
                if (wrap) dbg ? launch_sparse<2, true, true>(h, a, nb) : launch_sparse<2, true, false>(h, a, nb);
                else dbg ? launch_sparse<2, false, true>(h, a, nb) : launch_sparse<2, false, false>(h, a, nb);
            }
            LAUNCH_CHECK();
            continue;
        }
        if (h->dim == 3) {
            if (wrap) dbg ? launch_force<3, true, true>(h, a, nb, bs, lds) : launch_force<3, true, false>(h, a, nb, bs, lds);
            else dbg ? launch_force<3, false, true>(h, a, nb, bs, lds) : launch_force<3, false, false>(h, a, nb, bs, lds);
        } else {
            if (wrap) dbg ? launch_force<2, true, true>(h, a, nb, bs, lds) : launch_force<2, true, false>(h, a, nb, bs, lds);
            else dbg ? launch_force<2, false, true>(h, a, nb, bs, lds) : launch_force<2, false, false>(h, a, nb, bs, lds);
        }
        LAUNCH_CHECK();
    }
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

// ---------------------------------------------------------------------------
// launch planning and the step
// ---------------------------------------------------------------------------

// Launch shape of the force kernel for the current batch.  Explicit tuning
// fields win.  Otherwise the planner reads the particle-weighted cell
// occupancy (sum over cells of count^2, divided by N: the mean size of the
// cell a particle sits in) measured by an earlier step:
//   * uniform worlds: 256 homes per workgroup and a staging buffer that holds
//     the whole neighbourhood (one window); 24-entry candidate lists where
//     cells are dense enough for the LDS to limit occupancy, 32 where sparse;
//   * clustered 3D worlds (occupancy far above the mean N / cells): 128 homes
//     per workgroup and a small staging buffer, so that many workgroups share
//     a CU while the dense neighbourhoods stream through it in windows.
// The plan never changes results: every home visits its candidates in the
// same order whatever the windows and workgroup size.
ForcePlan plan_force_for(int dim, double avg, double occ, bool clustered, const plife_tuning &t,
                         bool dbg, int ntypes_max)
{
    const int NR = dim == 3 ? 9 : 3;
    ForcePlan p;
    p.bs = t.block_homes ? t.block_homes : (clustered ? 128 : 256);
    p.lcap = t.list_cap > 0 ? std::max(8, std::min(t.list_cap, 512))   // >= the screen group (8)
                            : (clustered || occ < 12.0 ? 32 : 24);
    p.lcap = std::min(p.lcap, 16384 / p.bs);             // lists stay within 32 KB of LDS
    // a workgroup of bs homes spans ~bs/avg + 1 cells of a row; it stages NR
    // rows of (span + 2) cells.  Clustered worlds keep a short cell table: the
    // few sparse workgroups that span more cells take the global path.
    const double span = p.bs / std::max(avg, 1e-9) + 1.0;
    // balanced path (uniform worlds): a home's stencil row is a run of three
    // cells, ~3 avg candidates, ceil(run / 32) mask words; the workgroup's
    // pool holds 10% more than the expected total (workgroups needing more
    // take the windowed path)
    if (!clustered && !t.windowed_only) {
        const double run = 3.0 * avg;
        const long long pool = (long long)std::ceil(p.bs * NR * (run / 32.0 + 0.5) * 1.1) + 64;
        if (pool <= 65535 && (size_t)pool * 6 <= kWordLds) p.wpool = (int)pool;
    }
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
    const bool m_in_lds = ntypes_max <= kMaxTypesLds;
    auto total = [&](long long c) {
        return force_lds_layout(dim, dbg, ntypes_max, m_in_lds, (int)c, p.lcap, p.bs,
                                p.max_cells, p.wpool).total;
    };
    // keep room for a useful staging window: shrink the cell table first
    const long long min_cap = std::min<long long>(want, 256);
    while (p.max_cells > 64 && total(min_cap) > budget) p.max_cells = std::max(64, p.max_cells * 3 / 4);
    // then shrink the staging capacity until the layout fits the LDS budget
    while (want > 64 && total(want) > budget) want -= 64;
    p.cap = (int)std::max(0LL, want);
    p.lds = total(p.cap);
    return p;
}

// The tuning the planner follows: the caller's, with the measured homes per
// workgroup where block_homes is automatic.
plife_tuning plan_tuning(const plife_physics *h)
{
    plife_tuning t = h->tune;
    if (t.block_homes == 0 && h->auto_bs != 0 && h->bs_meas == h->version) t.block_homes = h->auto_bs;
    return t;
}

ForcePlan plan_force(const plife_physics *h)
{
    const int dim = h->dim == 3 ? 3 : 2;
    // density of the worlds the staged kernel steps (all of them if none is)
    long long sn = 0, sc = 0;
    for (const WorldHost &w : h->worlds)
        if (!w.sparse) {
            sn += w.n;
            sc += grid_of(w.st).ncells;
        }
    if (sc == 0) {
        sn = h->n;
        sc = h->ncells;
    }
    const double avg = (double)sn / (double)std::max(sc, 1LL);
    const double occ = h->occupancy > 0.0 ? h->occupancy : avg + 1.0;   // Poisson: avg + 1
    return plan_force_for(dim, avg, occ, h->clustered, plan_tuning(h), h->debug != 0, h->ntypes_max);
}

// The class of a layout for the measured homes per workgroup: dimension,
// clustering, batch or single, the staged worlds' mean particles per cell in
// half octaves and the particle count in octaves.
uint64_t shape_class(const plife_physics *h)
{
    long long sn = 0, sc = 0;
    for (const WorldHost &w : h->worlds)
        if (!w.sparse) {
            sn += w.n;
            sc += grid_of(w.st).ncells;
        }
    const double avg = sc > 0 ? (double)sn / (double)sc : 0.0;
    const int da = (int)std::floor(std::log2(std::max(avg, 1.0 / 64.0)) * 2.0) + 64;
    const int nn = (int)std::floor(std::log2((double)std::max(h->n, 1)));
    return (uint64_t)(h->dim == 3) | (uint64_t)h->clustered << 1 | (uint64_t)(h->batch != 0) << 2 |
           (uint64_t)(uint32_t)da << 8 | (uint64_t)(uint32_t)nn << 24;
}

// Whether the next long update measures the homes per workgroup: the
// multi-launch path with block_homes automatic, some world on the staged
// force kernel, and no measurement for this layout or its class yet.
bool shape_tune_due(const plife_physics *h)
{
    if (h->small || h->tune.block_homes != 0 || h->tune.small_step == 2 || h->n <= 0 ||
        h->bs_meas == h->version)
        return false;
    for (const WorldHost &w : h->worlds)
        if (!w.sparse && w.n > 0) return true;
    return false;
}

// Read the clustering measure of the latest step the host has seen complete
// (a pinned copy refreshed by every step; a stale value only affects speed).
void refresh_occupancy(plife_physics *h)
{
    const int W = (int)h->worlds.size();
    // per-world neighbour counts, when the copy comes from the current layout
    if (h->stat_host && (size_t)(3 + W) <= h->stat_cap &&
        __atomic_load_n(h->stat_host + 2 + W, __ATOMIC_RELAXED) == h->version + 1)
        for (int w = 0; w < W; ++w) {
            const uint64_t p = __atomic_load_n(h->stat_host + 2 + w, __ATOMIC_RELAXED);
            if (h->worlds[w].n > 0 && h->worlds[w].steps > 0)
                h->worlds[w].nb = (double)p / (double)h->worlds[w].n;
        }
    const uint64_t q = h->stat_host ? __atomic_load_n(h->stat_host + 1, __ATOMIC_RELAXED) : 0;
    if (q == 0 || h->n <= 0 || h->ncells <= 0) return;
    h->occupancy = (double)q / (double)h->n;
    const double avg = (double)h->n / (double)h->ncells;
    // clustered: cells seen by a particle hold far more than the average (a
    // Poisson world has occupancy avg + 1) and enough to fill long windows
    // (3D only: in 2D the uniform plan measured as fast or faster on clustered
    // worlds and batches, e.g. C5's 2D jobs after 600 steps 18.0 vs 19.5 ms)
    const double ratio = h->occupancy / (avg + 1.0);
    if (h->dim != 3) h->clustered = false;
    else if (!h->clustered && ratio > 4.0 && h->occupancy > 64.0) h->clustered = true;
    else if (h->clustered && (ratio < 3.0 || h->occupancy < 48.0)) h->clustered = false;
}

// Table position of the block dispatched as blockIdx b of a launch of nb
// blocks (the kernels' xcd_remap).
inline int remap_pos(int b, int nb)
{
    const int q = nb >> 3, rem = nb & 7, xcd = b & 7, k = b >> 3;
    return xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
}

// Orders each segment of the block table heaviest first when the work of its
// blocks (their pairs in the last measured step) is skewed and a heavy block
// is dispatched late: workgroups start in dispatch order, so a heavy block
// near the end keeps a few CUs busy after the rest have drained.  Blocks are
// dealt in decreasing work round-robin to the 8 XCDs (blockIdx i -> table
// position remap_pos(i)).  Results do not depend on the order.
int balance_blocks(plife_physics *h)
{
    const int nb = h->nb_total;
    if (nb < 64 || (int)h->block_est.size() != nb) return PLIFE_OK;
    // the last step's measurement of this table, else the estimates it was built with
    std::vector<uint32_t> work(h->block_est);
    if (h->work_host && h->work_host_cap > (size_t)nb &&
        __atomic_load_n(h->work_host + nb, __ATOMIC_RELAXED) == (uint32_t)(h->blocks_serial + 1))
        for (int i = 0; i < nb; ++i) work[i] = __atomic_load_n(h->work_host + i, __ATOMIC_RELAXED);
    bool changed = false;
    int base = 0;
    for (int s = 0; s < 4; ++s) {
        const int n = h->seg[s];
        const int b0 = base;
        base += n;
        if (n < 64) continue;
        double mean = 0.0;
        uint32_t late = 0;              // heaviest block dispatched in the second half
        for (int b = 0; b < n; ++b) {
            const uint32_t wk = work[b0 + remap_pos(b, n)];
            mean += wk;
            if (b >= n / 2) late = std::max(late, wk);
        }
        mean /= n;
        if (!(late > 2.0 * mean) || !(mean > 0.0)) continue;
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return work[b0 + x] > work[b0 + y]; });
        std::vector<uint2> blk(n);
        std::vector<uint64_t> key(n);
        std::vector<uint32_t> est(n);
        for (int i = 0; i < n; ++i) {
            const int to = remap_pos(i, n), from = b0 + order[i];
            blk[to] = h->blocks[from];
            key[to] = h->block_key[from];
            est[to] = work[from];
        }
        std::copy(blk.begin(), blk.end(), h->blocks.begin() + b0);
        std::copy(key.begin(), key.end(), h->block_key.begin() + b0);
        std::copy(est.begin(), est.end(), h->block_est.begin() + b0);
        changed = true;
    }
    if (!changed) return PLIFE_OK;
    int rc;
    const size_t bytes = h->blocks.size() * sizeof(uint2);
    if ((rc = begin_uploads(h, bytes))) return rc;
    if ((rc = upload(h, h->d_blocks, h->blocks.data(), bytes))) return rc;
    if ((rc = end_uploads(h))) return rc;
    ++h->blocks_serial;
    return PLIFE_OK;
}

// Which force kernel steps each world.  A world goes to the sparse-world
// kernel when a staged workgroup of its homes would span more than two grid
// rows or more cells than the staged cell table holds; the span follows from
// the occupancy of the cells its particles sit in, estimated from the
// neighbours per particle of the last measured step (a uniform world's count
// before that): local density nb / V(rmax) times the cell volume rmax^dim.
// Grids of fewer than 3 cells per axis always take it (literal stencil over
// repeated cells).  Hysteresis keeps worlds near the threshold from
// switching kernels, and so recapturing, at every update.  Results do not
// depend on the class.
void classify_worlds(plife_physics *h)
{
    const int dim = h->dim == 3 ? 3 : 2;
    const int NR = dim == 3 ? 9 : 3;
    const int mode = h->tune.sparse_worlds;
    const plife_tuning pt = plan_tuning(h);
    const double bs = pt.block_homes ? pt.block_homes : (h->clustered ? 128 : 256);
    bool changed = false;
    for (WorldHost &w : h->worlds) {
        const Grid gr = grid_of(w.st);
        bool sparse;
        if (gr.nx < 3) {
            sparse = true;
        } else if (mode != 0) {
            sparse = mode == 2;
        } else {
            const double r = w.st.rmax;
            const double cell = dim == 3 ? r * r * r : r * r;
            const double ball = dim == 3 ? 4.0 / 3.0 * M_PI * r * r * r : M_PI * r * r;
            const double nb = w.nb >= 0.0 ? w.nb : std::max(w.n - 1, 0) * ball;
            const double avg = (double)w.n / (double)gr.ncells;
            const double occ = std::max({avg, nb * cell / ball, 1e-9});
            const double span = bs / occ;              // cells spanned by a workgroup's homes
            const double k = w.sparse ? 0.75 : 1.0;
            sparse = span > 2.0 * gr.nx * k || NR * (span + 6.0) > 2048.0 * k;
        }
        if (sparse != w.sparse) {
            w.sparse = sparse;
            changed = true;
        }
    }
    if (changed) ++h->classes;
}

// The one-launch step's kernel for the handle's dimension, debug mode and
// matrix placement.
template <int DIM, bool DEBUG, bool MLDS>
const void *small_kernel_of()
{
    void (*f)(SmallArgs) = k_step_small<DIM, DEBUG, MLDS>;
    return reinterpret_cast<const void *>(f);
}

const void *small_kernel(int dim, bool dbg, bool mlds)
{
    if (dim == 3) {
        if (dbg) return mlds ? small_kernel_of<3, true, true>() : small_kernel_of<3, true, false>();
        return mlds ? small_kernel_of<3, false, true>() : small_kernel_of<3, false, false>();
    }
    if (dbg) return mlds ? small_kernel_of<2, true, true>() : small_kernel_of<2, true, false>();
    return mlds ? small_kernel_of<2, false, true>() : small_kernel_of<2, false, false>();
}

// Whether the handle steps in one launch (k_step_small): every world fits one
// workgroup's LDS (particles, cells, types), and the tuning allows it.  Homes
// per workgroup: the fewest of 256, 512, 1024, ... for which the workgroups of
// all worlds fit one pass over the CUs (each workgroup sorts its whole world,
// so fewer, larger workgroups waste less once the device is full); when none
// does, one workgroup per world.
bool plan_small(plife_physics *h)
{
    const plife_tuning &t = h->tune;
    if (t.small_step == 1 || h->n <= 0 || h->worlds.empty()) return false;
    if (t.small_step == 0 && (t.block_homes || t.list_cap || t.stage_slack > 0.0 || t.sparse_worlds ||
                              t.windowed_only))
        return false;
    const int dim = h->dim == 3 ? 3 : 2;
    int nmax = 0;
    long long cmax = 0;
    for (const WorldHost &w : h->worlds) {
        const long long nc = grid_of(w.st).ncells;
        if (w.n > kSmallMaxN || nc > kSmallMaxCells) return false;
        nmax = std::max(nmax, w.n);
        cmax = std::max(cmax, nc);
    }
    const bool mlds = h->ntypes_max <= kMaxTypesLds;
    const size_t lds = small_lds_layout(dim, mlds ? h->ntypes_max : 0, nmax, (int)cmax).total;
    if (lds + 1024 > (size_t)h->lds_limit) return false;     // + the kernel's static LDS
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(small_kernel(dim, h->debug != 0, mlds),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
    }
    int homes = 256;
    if (t.small_step == 2 && t.block_homes) {
        homes = t.block_homes;
    } else {
        int cus = 256;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess)
            cus = 256;
        for (homes = 128;; homes *= 2) {
            long long blocks = 0;
            for (const WorldHost &w : h->worlds) blocks += std::max(1, ceil_div(w.n, homes));
            if (blocks <= cus || homes >= nmax) break;
        }
        homes = std::max(128, std::min(homes, ((nmax + 127) / 128) * 128));
    }
    h->small_homes = homes;
    h->small_nmax = nmax;
    h->small_cmax = (int)cmax;
    h->small_lds = lds;
    return true;
}

// The one-launch step's workgroup table: (world, first home slot) for each
// run of small_homes homes of every world (one entry for an empty world, which
// still writes its cell table).
int ensure_small_blocks(plife_physics *h)
{
    if (h->blocks_bs == -h->small_homes && h->blocks_version == h->version &&
        h->blocks_classes == h->classes)
        return PLIFE_OK;
    h->blocks.clear();
    h->block_key.clear();
    h->block_est.clear();
    for (int w = 0; w < (int)h->worlds.size(); ++w) {
        const WorldHost &wh = h->worlds[w];
        for (int g = 0; g == 0 || g < wh.n; g += h->small_homes) {
            h->blocks.push_back(make_uint2((uint32_t)w, (uint32_t)(wh.off + g)));
            h->block_key.push_back(~0ull);
            h->block_est.push_back(0u);
        }
    }
    for (int &sg : h->seg) sg = 0;
    h->nb_total = (int)h->blocks.size();
    int rc;
    if ((rc = ensure_dev(&h->d_blocks, &h->blocks_cap, h->blocks.size(), &h->retired))) return rc;
    ++h->blocks_serial;
    const size_t bytes = h->blocks.size() * sizeof(uint2);
    if ((rc = begin_uploads(h, bytes))) return rc;
    if ((rc = upload(h, h->d_blocks, h->blocks.data(), bytes))) return rc;
    if ((rc = end_uploads(h))) return rc;
    h->blocks_bs = -h->small_homes;
    h->blocks_version = h->version;
    h->blocks_classes = h->classes;
    return PLIFE_OK;
}

// Everything a step of the current configuration needs, so that the steps
// (captured or direct) upload and allocate nothing.
int prepare_step(plife_physics *h)
{
    int rc;
    // a matrix made smaller than the particles' types (plife_set_settings) needs
    // particles within it before the next step: the force law indexes the matrix by type
    for (const WorldHost &w : h->worlds)
        if (w.live && w.n > 0 && w.max_type >= w.ntypes)
            return fail(PLIFE_ERR_STATE,
                        "world %d: particle types reach %d but the matrix has %d types; set particles "
                        "with types below the matrix size before stepping", w.id, w.max_type, w.ntypes);
    if ((rc = sync_layout(h))) return rc;
    refresh_occupancy(h);
    classify_worlds(h);
    // a layout of a class measured before takes that class's homes per workgroup
    if (h->tune.block_homes == 0 && h->bs_meas != h->version) {
        const uint64_t c = shape_class(h);
        for (const auto &e : h->bs_seen)
            if (e.first == c) {
                h->auto_bs = e.second;
                h->bs_meas = h->version;
                classify_worlds(h);
            }
    }
    h->plan = plan_force(h);
    if ((rc = ensure_cells(h, h->ncells))) return rc;
    if ((rc = ensure_debug(h))) return rc;
    // the one-launch step where it fits: forced by the tuning, or chosen by
    // timing both paths on this layout (run_steps); the multi-launch path's
    // buffers are prepared either way
    h->small_ok = plan_small(h);
    h->small = h->small_ok && (h->tune.small_step == 2 || (h->small_meas == h->version && h->small_pref));
    const RadixPlan rp = radix_plan(h->ncells, h->n);
    if ((rc = ensure_dev(&h->hist, &h->hist_cap, rp.hist_len, &h->retired))) return rc;
    const size_t nb = (size_t)ceil_div(h->ncells, kScanTile) +
                      (size_t)ceil_div((long long)rp.hist_len, kScanTile);
    if ((rc = ensure_dev(&h->bsum, &h->bsum_cap, nb, &h->retired))) return rc;
    if ((rc = ensure_dev(&h->bsq, &h->bsq_cap, (size_t)ceil_div(h->ncells, kScanTile), &h->retired)))
        return rc;
    if (h->small) return ensure_small_blocks(h);
    if ((rc = ensure_blocks(h, h->plan.bs))) return rc;
    return balance_blocks(h);
}

// Switches the prepared step between the one-launch and the multi-launch path.
int select_path(plife_physics *h, bool one)
{
    h->small = one;
    if (one) return ensure_small_blocks(h);
    int rc = ensure_blocks(h, h->plan.bs);
    return rc ? rc : balance_blocks(h);
}

// One step of every world in one launch of k_step_small.
int do_small_step(plife_physics *h, PhaseEvents *ev)
{
    int rc;
    if (h->blocks_bs != -h->small_homes || h->blocks_version != h->version ||
        h->blocks_classes != h->classes || h->layout_dirty || h->params_dirty)
        return fail(PLIFE_ERR_STATE, "the step was not prepared");
    for (int k = 0; k < 4; ++k)
        if ((rc = record(h, ev, k))) return rc;
    const int pi = h->pc, ai = h->ac, po = 1 - h->pc, ao = 1 - h->ac;
    SmallArgs s{};
    ForceArgs &a = s.f;
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
    a.matrix = h->d_matrix;
    a.worlds = h->d_worlds;
    a.blocks = h->d_blocks;
    a.ntypes_max = h->ntypes_max;
    a.m_in_lds = h->ntypes_max <= kMaxTypesLds ? 1 : 0;
    a.dbg_count = h->dbg_count;
    a.dbg_hash = h->dbg_hash;
    a.dbg_acc = h->dbg_acc;
    s.perm_out = h->vA;
    // the last step's sorted keys: the incremental sort's hint (checked in the kernel)
    s.skeys_in = h->small_kbuf ? h->kA : h->skeys;
    s.skeys_out = h->small_kbuf ? h->skeys : h->kA;
    s.keys_out = h->keys0;
    s.cell_start_out = h->cell_start;
    s.cell_end_out = h->cell_end;
    s.acc = h->d_facc;
    s.counters = h->d_pairs;
    s.stat_out = h->stat_dev;
    s.tag = h->version + 1;
    s.ncounters = 2 + (int)h->worlds.size();
    s.homes = h->small_homes;
    s.n_max = h->small_nmax;
    s.cells_max = h->small_cmax;
    const bool dbg = h->debug != 0;
    const int nb = h->nb_total;
    const size_t lds = h->small_lds;
    if (h->dim == 3) {
        if (dbg) a.m_in_lds ? k_step_small<3, true, true><<<nb, kSmallThreads, lds, h->stream>>>(s)
                            : k_step_small<3, true, false><<<nb, kSmallThreads, lds, h->stream>>>(s);
        else a.m_in_lds ? k_step_small<3, false, true><<<nb, kSmallThreads, lds, h->stream>>>(s)
                        : k_step_small<3, false, false><<<nb, kSmallThreads, lds, h->stream>>>(s);
    } else {
        if (dbg) a.m_in_lds ? k_step_small<2, true, true><<<nb, kSmallThreads, lds, h->stream>>>(s)
                            : k_step_small<2, true, false><<<nb, kSmallThreads, lds, h->stream>>>(s);
        else a.m_in_lds ? k_step_small<2, false, true><<<nb, kSmallThreads, lds, h->stream>>>(s)
                        : k_step_small<2, false, false><<<nb, kSmallThreads, lds, h->stream>>>(s);
    }
    LAUNCH_CHECK();
    h->stepped = true;
    h->last_ncells = h->ncells;
    h->perm = h->vA;
    h->small_kbuf ^= 1;
    h->last_lds = lds;
    h->last_bs = h->small_homes;
    if ((rc = record(h, ev, 4))) return rc;
    h->pc = po;
    h->ac = ao;
    return PLIFE_OK;
}

// One step of every world.  prepare_step must have run since the last change.
int do_step(plife_physics *h, PhaseEvents *ev)
{
    if (h->small) return do_small_step(h, ev);
    const int n = h->n;
    const int nc = (int)h->ncells;
    int rc;
    if (h->blocks_bs != h->plan.bs || h->blocks_version != h->version ||
        h->blocks_classes != h->classes || h->layout_dirty ||
        h->params_dirty)
        return fail(PLIFE_ERR_STATE, "the step was not prepared");

    if ((rc = record(h, ev, 0))) return rc;
    const RadixPlan plan = radix_plan(h->ncells, n);
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
        if (h->dim == 3) launch_bin<3>(h, plan.D, P0, n, plan.ntiles);
        else launch_bin<2>(h, plan.D, P0, n, plan.ntiles);
    } else {
        const int nz = 2 + (int)h->worlds.size();
        k_zero_u64<<<ceil_div(nz, kBlock), kBlock, 0, h->stream>>>(h->d_pairs, nz);
    }
    LAUNCH_CHECK();
    if ((rc = scan_jobs(h, ScanJob{h->cell_count, nc, h->cell_start, h->cell_end, 1, nullptr,
                                   h->d_pairs + 1},
                        n > 0 ? ScanJob{h->hist, (int)plan.hist_len, h->hist, nullptr, 0, nullptr, nullptr}
                              : no_job())))
        return rc;
    h->cells_dirty = false;
    if ((rc = record(h, ev, 1))) return rc;

    h->stepped = true;
    h->last_ncells = h->ncells;
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

    // 3. force + integrate (Physics.java:122-131 fused: the update is Jacobi):
    //    reads the pre-sort state through the permutation and writes the new
    //    state in cell order (the array swap, Physics.java:350-353)
    const int pi = h->pc, ai = h->ac, po = 1 - h->pc, ao = 1 - h->ac;
    const ForcePlan &fp = h->plan;
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
    a.skeys = h->skeys;
    a.cell_start = h->cell_start;
    a.cell_end = h->cell_end;
    a.matrix = h->d_matrix;
    a.worlds = h->d_worlds;
    a.blocks = h->d_blocks;
    a.ntypes_max = h->ntypes_max;
    a.m_in_lds = h->ntypes_max <= kMaxTypesLds ? 1 : 0;
    a.cap = fp.cap;
    a.lcap = fp.lcap;
    a.max_cells = fp.max_cells;
    a.wpool = fp.wpool;
    a.pair_total = h->d_pairs;
    a.pair_world = h->d_pairs + 2;
    a.block_work = h->work_dev;
    a.work_tag = (uint32_t)(h->blocks_serial + 1);
    a.nb_total = h->nb_total;
    a.dbg_count = h->dbg_count;
    a.dbg_hash = h->dbg_hash;
    a.dbg_acc = h->dbg_acc;
    h->last_lds = fp.lds;
    h->last_cap = fp.cap;
    h->last_bs = fp.bs;
    h->last_lcap = fp.lcap;
    if ((rc = launch_force_any(h, a, fp.bs, fp.lds))) return rc;
    h->small_kbuf = 0;                 // the last radix pass wrote the sorted keys to skeys
    if ((rc = record(h, ev, 4))) return rc;
    h->pc = po;
    h->ac = ao;
    return PLIFE_OK;
}

std::vector<unsigned char> graph_key(const plife_physics *h)
{
    struct Key {
        int n, debug, pc, ac, dim, pad;
        long long ncells;
        uint64_t version;
        plife_tuning tune;
        ForcePlan plan;
        int seg[4];
        int small, small_homes, small_kbuf, pad2;
        uint64_t blocks_serial;
        const void *ptr[40];
    } k;
    std::memset(&k, 0, sizeof k);
    k.n = h->n;
    k.debug = h->debug;
    k.pc = h->pc;
    k.ac = h->ac;
    k.dim = h->dim;
    k.ncells = h->ncells;
    k.version = h->version;
    k.tune = h->tune;
    k.plan = h->plan;
    k.small = h->small ? 1 : 0;
    k.small_homes = h->small_homes;
    k.small_kbuf = h->small_kbuf;
    std::memcpy(k.seg, h->seg, sizeof k.seg);
    k.blocks_serial = h->blocks_serial;
    const void *p[] = {h->px[0], h->px[1], h->py[0], h->py[1], h->pz[0], h->pz[1], h->vx[0],
                       h->vx[1], h->vy[0], h->vy[1], h->vz[0], h->vz[1], h->type[0], h->type[1],
                       h->id[0], h->id[1], h->keys0, h->kA, h->kB, h->vA, h->vB, h->cell_count,
                       h->cell_start, h->cell_end, h->hist, h->bsum, h->bsq, h->dbg_count, h->dbg_hash, h->dbg_acc,
                       h->d_matrix, h->d_pairs, h->d_worlds, h->world_of, h->d_blocks,
                       h->stat_host, h->work_host, h->skeys, h->d_facc};
    std::memcpy(k.ptr, p, sizeof p);
    std::vector<unsigned char> v(sizeof k);
    std::memcpy(v.data(), &k, sizeof k);
    return v;
}

// Destroys the step graph.  A launch of it may still be executing on the
// stream (updates return before the device finishes), so the stream is drained
// first: the executable graph's resources are never freed under a running
// launch.
void drop_graph(plife_physics *h)
{
    if (h->gexec) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(h->gexec);
    }
    h->gexec = nullptr;
    h->gkey.clear();
}

// Captures two consecutive steps into h->gexec (not executed yet; two steps
// leave the buffer roles where they started).
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

void count_steps(plife_physics *h, int k)
{
    for (WorldHost &w : h->worlds) w.steps += k;
}

// After the last step of an update: its counters to the host's pinned copy
// for the launch planner (each step's bin kernel copies the previous step's).
int copy_stats(plife_physics *h)
{
    if (!h->stepped || !h->d_pairs || !h->stat_dev) return PLIFE_OK;
    const int nc = 3 + (int)h->worlds.size();
    k_copy_u64<<<ceil_div(nc, kBlock), kBlock, 0, h->stream>>>(h->d_pairs, h->stat_dev, nc);
    LAUNCH_CHECK();
    return PLIFE_OK;
}

// Steps plife_update keeps queued on the device.  The reference's stop request
// takes effect before the next particle (Physics.java:118-133, 149-151); a
// device step cannot stop halfway, so plife_update queues only about
// kStopWindowMs of work ahead of the device and checks the stop flag before
// every launch: a stop then leaves at most that much work to finish.  The
// step time is estimated from the last measured step (pairs and particles).
int flight_limit(const plife_physics *h)
{
    const double pairs = h->stat_host ? (double)__atomic_load_n(h->stat_host, __ATOMIC_RELAXED) : 0.0;
    const double est_ms = 0.03 + 4e-9 * pairs + 3e-7 * (double)h->n;
    const double lim = std::ceil(kStopWindowMs / est_ms);
    return (int)std::min<double>(std::max(lim, 2.0), (double)kFlightSlots - 2);
}

// Marks the end of a launch of k steps on the stream; waits for the oldest
// queued launches while more than `limit` steps are queued.
int throttle(plife_physics *h, int k, int limit)
{
    while (h->fl_count > 0 && (h->fl_queued + k > limit || h->fl_count == kFlightSlots)) {
        const int s = (h->fl_head - h->fl_count + kFlightSlots) % kFlightSlots;
        HIP_TRY(hipEventSynchronize(h->fl_ev[s]));
        h->fl_queued -= h->fl_steps[s];
        --h->fl_count;
    }
    HIP_TRY(hipEventRecord(h->fl_ev[h->fl_head], h->stream));
    h->fl_steps[h->fl_head] = k;
    h->fl_head = (h->fl_head + 1) % kFlightSlots;
    ++h->fl_count;
    h->fl_queued += k;
    return PLIFE_OK;
}

// nsteps steps: pairs of steps replay a captured graph, the rest launch
// directly.  A graph is captured once the same configuration comes back on a
// second update (or when one update runs many steps); a configuration that
// changes every update (worlds joining and leaving) launches directly, as a
// capture would not be replayed enough to pay for itself.  With honour_stop
// the stop flag is checked before every launch; a single-world handle then
// also bounds the queued work (flight_limit), so that the steps a stop leaves
// behind are few.  Batches enqueue every step at once: their updates are
// short quanta (the serving loop steps at most 32 at a time) and two batches
// on their own streams must overlap.
int run_steps(plife_physics *h, int nsteps, bool honour_stop)
{
    int k = 0, rc;
    if ((rc = prepare_step(h))) return rc;
    const bool bounded = honour_stop && !h->batch;
    const int limit = bounded ? flight_limit(h) : 0;
    h->fl_count = h->fl_queued = 0;
    auto launched = [&](int steps) -> int {
        count_steps(h, steps);
        return bounded ? throttle(h, steps, limit) : PLIFE_OK;
    };
    // The first update of a 3D handle: the launch plan follows the clustering
    // of the cells, which only a step measures, so a long first update steps
    // once, waits for that step's counters and re-plans the rest (C3: the
    // clustered plan is 1.6x faster than the uniform one it would otherwise run).
    if (h->dim == 3 && h->occupancy <= 0.0 && !h->small && nsteps - k >= 3 && h->n > 0 &&
        h->tune.block_homes == 0) {
        if (honour_stop && h->stop.load()) return PLIFE_STOPPED;
        if ((rc = do_step(h, nullptr))) return rc;
        ++k;
        if ((rc = launched(1))) return rc;
        if ((rc = copy_stats(h))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
        if ((rc = prepare_step(h))) return rc;
    }
    // Four steps of the prepared configuration timed as the update runs them
    // (two replays of the step graph, or direct launches under
    // direct_launch); the one-launch step first comes back to key buffer 0
    // for its graph.  Every configuration gives the same bits, so these are
    // steps of the update.
    const bool graphs = !h->tune.direct_launch;
    auto time_four = [&](float *ms) -> int {
        if (graphs && h->small && h->small_kbuf != 0) {
            if ((rc = do_step(h, nullptr))) return rc;
            k += 1;
            if ((rc = launched(1))) return rc;
        }
        if (graphs) {
            std::vector<unsigned char> key = graph_key(h);
            if (!h->gexec || key != h->gkey) {
                if ((rc = capture_two_steps(h))) return rc;
                h->gkey = key;
                h->seen_key = key;
            }
        }
        HIP_TRY(hipEventRecord(h->tune_ev[0], h->stream));
        if (graphs) {
            for (int i = 0; i < 2; ++i) HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
        } else {
            for (int i = 0; i < 4; ++i)
                if ((rc = do_step(h, nullptr))) return rc;
        }
        HIP_TRY(hipEventRecord(h->tune_ev[1], h->stream));
        HIP_TRY(hipEventSynchronize(h->tune_ev[1]));
        HIP_TRY(hipEventElapsedTime(ms, h->tune_ev[0], h->tune_ev[1]));
        k += 4;
        return launched(4);
    };
    // Two configurations, 0 (the default) and 1, timed in the order A B B A
    // (time_four each), so that a clock still rising at the start of the
    // update, or any drift, weighs on both alike; t[c] sums configuration c's
    // two timings.  Each switch steps the configuration once and prepares it
    // again, so that its workgroups are ordered by that step's measured work
    // as they are in the update's later steps.
    auto time_pair = [&](auto &&configure, float t[2]) -> int {
        static constexpr int order[4] = {0, 1, 1, 0};
        t[0] = t[1] = 0.0f;
        for (int i = 0; i < 4; ++i) {
            if (honour_stop && h->stop.load()) return PLIFE_STOPPED;
            if (i == 0 || order[i] != order[i - 1]) {
                if ((rc = configure(order[i]))) return rc;
                if ((rc = do_step(h, nullptr))) return rc;
                k += 1;
                if ((rc = launched(1))) return rc;
                if ((rc = copy_stats(h))) return rc;
                HIP_TRY(hipStreamSynchronize(h->stream));
                if ((rc = configure(order[i]))) return rc;
            }
            float ms = 0.0f;
            if ((rc = time_four(&ms))) return rc;
            t[order[i]] += ms;
        }
        return PLIFE_OK;
    };
    // the alternative replaces the default only when measurably faster
    auto faster = [](const float t[2]) { return t[1] < 0.97f * t[0]; };
    // Automatic tuning of a layout that fits the one-launch step: its first
    // update of at least kTuneSteps steps times the multi-launch path against
    // the one-launch step (time_pair) and keeps the faster for the layout.
    // The redundant per-workgroup sort makes the one-launch step faster for
    // small worlds and batches that leave CUs idle, slower once the device is
    // full.
    if (h->small_ok && h->tune.small_step == 0 && h->small_meas != h->version && nsteps >= kTuneSteps &&
        nsteps - k >= kTuneMax) {
        float t[2];
        rc = time_pair([&](int p) { return select_path(h, p == 1); }, t);
        if (rc == PLIFE_STOPPED && k > 0) (void)copy_stats(h);
        if (rc) return rc;
        h->small_pref = faster(t);
        h->small_meas = h->version;
        if ((rc = select_path(h, h->small_pref))) return rc;
    }
    // Automatic homes per force workgroup on the multi-launch path: the first
    // update of at least kTuneSteps steps of a class of layout (shape_class)
    // times the planner's choice (256, or 128 for clustered 3D) against the
    // other (time_pair) and keeps the faster for every layout of the class.
    if (shape_tune_due(h) && nsteps >= kTuneSteps && nsteps - k >= kTuneMax) {
        const int cand[2] = {h->clustered ? 128 : 256, h->clustered ? 256 : 128};   // the planner's first
        float t[2];
        rc = time_pair([&](int p) {
            h->auto_bs = cand[p];
            h->bs_meas = h->version;
            return prepare_step(h);
        }, t);
        if (rc) {
            // an interrupted measurement chooses nothing: the next long update measures again
            h->auto_bs = 0;
            h->bs_meas = ~0ull;
            if (rc == PLIFE_STOPPED && k > 0) (void)copy_stats(h);
            return rc;
        }
        h->auto_bs = faster(t) ? cand[1] : cand[0];
        h->bs_seen.emplace_back(shape_class(h), h->auto_bs);
        if ((rc = prepare_step(h))) return rc;
    }
    bool use_graph = nsteps - k >= 2 && !h->tune.direct_launch;
    // the one-launch step alternates two key buffers; its graphs start from
    // buffer 0, after one direct step when the last update ended on buffer 1
    if (use_graph && h->small && h->small_kbuf != 0 && nsteps - k >= 3) {
        if (honour_stop && h->stop.load()) return PLIFE_STOPPED;
        if ((rc = do_step(h, nullptr))) return rc;
        ++k;
        if ((rc = launched(1))) return rc;
    }
    if (use_graph) {
        std::vector<unsigned char> key = graph_key(h);
        if (!h->gexec || key != h->gkey) {
            if (key != h->seen_key && nsteps - k < 64) use_graph = false;
            h->seen_key = key;
        }
    }
    if (use_graph) {
        std::vector<unsigned char> key = graph_key(h);
        if (!h->gexec || key != h->gkey) {
            if ((rc = capture_two_steps(h))) return rc;
            h->gkey = key;
        }
        while (nsteps - k >= 2) {
            if (honour_stop && h->stop.load()) {
                if (k > 0) (void)copy_stats(h);
                return PLIFE_STOPPED;
            }
            HIP_TRY(hipGraphLaunch(h->gexec, h->stream));
            k += 2;
            if ((rc = launched(2))) return rc;
        }
    }
    for (; k < nsteps; ++k) {
        if (honour_stop && h->stop.load()) {
            if (k > 0) (void)copy_stats(h);
            return PLIFE_STOPPED;
        }
        if ((rc = do_step(h, nullptr))) return rc;
        if ((rc = launched(1))) return rc;
    }
    return k > 0 ? copy_stats(h) : PLIFE_OK;
}

int time_steps(plife_physics *h, int nsteps, double *ms)
{
    int rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // build (or reuse) the step graph outside the timed region: the first
    // update of a new configuration launches directly, the second captures
    if ((rc = prepare_step(h))) return rc;
    // the automatic path choice of this layout, also outside the timed region
    if (h->small_ok && h->tune.small_step == 0 && h->small_meas != h->version && nsteps >= kTuneSteps) {
        if ((rc = run_steps(h, kTuneSteps, false))) return rc;
        if ((rc = prepare_step(h))) return rc;
    }
    if (shape_tune_due(h) && nsteps >= kTuneSteps) {
        if ((rc = run_steps(h, kTuneSteps, false))) return rc;
        if ((rc = prepare_step(h))) return rc;
    }
    for (int i = 0; i < 2 && nsteps >= 2 && !h->tune.direct_launch &&
                    (!h->gexec || graph_key(h) != h->gkey); ++i)
        if ((rc = run_steps(h, 2, false))) return rc;
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

int pair_total(plife_physics *h, int64_t *pairs)
{
    *pairs = 0;
    if (!h->d_pairs) return PLIFE_OK;
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, h->d_pairs, 8, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *pairs = (int64_t)v;
    return PLIFE_OK;
}

int profile(plife_physics *h, int nsteps, plife_timing *out)
{
    int rc;
    if ((rc = prepare_step(h))) return rc;
    std::vector<PhaseEvents> ev(nsteps);
    for (auto &pe : ev)
        for (auto &e : pe.e) HIP_TRY(hipEventCreate(&e));
    for (int k = 0; k < nsteps && !rc; ++k) {
        rc = do_step(h, &ev[k]);
        if (!rc) count_steps(h, 1);
    }
    if (!rc && nsteps > 0) rc = copy_stats(h);
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
        rc = pair_total(h, &pairs);
        out->pairs = pairs;
    }
    for (auto &pe : ev)
        for (auto &e : pe.e) (void)hipEventDestroy(e);
    return rc;
}

int check_settings(const plife_settings *s, int32_t ntypes, const double *matrix)
{
    if (!s || !matrix) return fail(PLIFE_ERR_ARGUMENT, "null argument");
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
    return PLIFE_OK;
}

int check_particles(int32_t n, const double *pos, const double *vel, const int32_t *types, int ntypes)
{
    if (n < 0) return fail(PLIFE_ERR_ARGUMENT, "n must be >= 0, got %d", n);
    if (n > 0 && (!pos || !vel || !types)) return fail(PLIFE_ERR_ARGUMENT, "null array");
    for (int i = 0; i < n; ++i)
        if (types[i] < 0 || types[i] >= ntypes)
            return fail(PLIFE_ERR_ARGUMENT, "particle %d has type %d outside [0, %d)", i, types[i], ntypes);
    return PLIFE_OK;
}

// Stores new settings of a world; marks the device table stale only on change.
void store_settings(plife_physics *h, WorldHost &w, const plife_settings *s, int32_t ntypes,
                    const double *matrix)
{
    const size_t mm = (size_t)ntypes * ntypes;
    const bool changed = std::memcmp(&w.st, s, sizeof *s) != 0 || ntypes != w.ntypes ||
                         w.matrix.size() != mm ||
                         std::memcmp(w.matrix.data(), matrix, mm * sizeof(double)) != 0;
    if (!changed) return;
    w.st = *s;
    w.ntypes = ntypes;
    w.matrix.assign(matrix, matrix + mm);
    h->params_dirty = true;
}

// Host threads for conversions: the CPUs this process may run on, at most 16.
int host_threads()
{
    cpu_set_t s;
    int n = 0;
    if (sched_getaffinity(0, sizeof(s), &s) == 0) n = CPU_COUNT(&s);
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 16));
}

// Runs f(lo, hi) over [0, n) on the host threads (one below 65,536 items).
template <typename F>
void parallel_for(int n, F f)
{
    const int nt = n >= kDlMinChunk ? host_threads() : 1;
    std::vector<std::thread> ts;
    for (int t = 1; t < nt; ++t)
        ts.emplace_back([&, t]() {
            const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
            if (lo < hi) f(lo, hi);
        });
    const int hi0 = (int)((long long)n / nt);
    if (hi0 > 0) f(0, hi0);
    for (auto &th : ts) th.join();
}

// The handle's pinned staging for up- and downloads, at least `bytes`.
int ensure_pin(plife_physics *h, size_t bytes)
{
    if (bytes <= h->pin_cap) return PLIFE_OK;
    if (h->pin) (void)hipHostFree(h->pin);
    h->pin = nullptr;
    h->pin_cap = 0;
    const hipError_t e = hipHostMalloc((void **)&h->pin, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        h->pin = nullptr;
        return fail(PLIFE_ERR_MEMORY, "hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    h->pin_cap = bytes;
    return PLIFE_OK;
}

// Converts n particles (Physics.ensurePosition, fixed point; FP32 velocities)
// on the host threads into pinned staging, moves them with one DMA and
// scatters them with k_unpack into slots [off, off + n) of the current
// buffers, with ids 0..n-1.  The caller has synchronised the stream.
int upload_particles(plife_physics *h, int off, int n, const double *pos, const double *vel,
                     const int32_t *types, bool wrap, bool d3)
{
    if (n == 0) return PLIFE_OK;
    const int na = d3 ? 7 : 5;           // x, y, [z], vx, vy, [vz], type
    const size_t words = (size_t)na * (size_t)n, nn = (size_t)n;
    int rc = ensure_pin(h, words * 4);
    if (rc) return rc;
    if ((rc = ensure_dev(&h->dl_dev, &h->dl_cap, words))) return rc;
    uint32_t *const st = (uint32_t *)h->pin;
    uint32_t *const sv = st + (d3 ? 3 : 2) * nn;
    parallel_for(n, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            st[i] = fixed_of(pos[3 * (size_t)i], wrap);
            st[nn + i] = fixed_of(pos[3 * (size_t)i + 1], wrap);
            if (d3) st[2 * nn + i] = fixed_of(pos[3 * (size_t)i + 2], wrap);
            const float fx = (float)vel[3 * (size_t)i], fy = (float)vel[3 * (size_t)i + 1];
            std::memcpy(sv + i, &fx, 4);
            std::memcpy(sv + nn + i, &fy, 4);
            if (d3) {
                const float fz = (float)vel[3 * (size_t)i + 2];
                std::memcpy(sv + 2 * nn + i, &fz, 4);
            }
        }
        std::memcpy(sv + (d3 ? 3 : 2) * nn + lo, types + lo, (size_t)(hi - lo) * 4);
    });
    HIP_TRY(hipMemcpyAsync(h->dl_dev, st, words * 4, hipMemcpyHostToDevice, h->stream));
    const int p = h->pc, a = h->ac;
    UnpackArgs u{};
    u.px = h->px[p];
    u.py = h->py[p];
    u.pz = h->pz[p];
    u.vx = h->vx[a];
    u.vy = h->vy[a];
    u.vz = h->vz[a];
    u.type = h->type[a];
    u.id = h->id[a];
    u.off = off;
    u.n = n;
    u.dim = d3 ? 3 : 2;
    k_unpack<<<ceil_div(n, kBlock), kBlock, 0, h->stream>>>(u, h->dl_dev);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PLIFE_OK;
}

// Generates n particles of the default presets from `seed` into slots
// [off, off + n) of the current buffers (ids 0..n-1).
int generate_particles(plife_physics *h, int off, int n, uint64_t seed, int ntypes, int dim)
{
    if (n == 0) return PLIFE_OK;
    GenArgs g{};
    for (int s = 0; s < 4; ++s) g.base[s] = splitmix_mix(seed + (uint64_t)(s + 1) * kStream);
    g.n = n;
    g.off = off;
    g.ntypes = ntypes;
    g.dim = dim;
    const int p = h->pc, a = h->ac;
    g.px = h->px[p];
    g.py = h->py[p];
    g.pz = h->pz[p];
    g.vx = h->vx[a];
    g.vy = h->vy[a];
    g.vz = h->vz[a];
    g.type = h->type[a];
    g.id = h->id[a];
    k_gen_uniform<<<ceil_div(n, kBlock), kBlock, 0, h->stream>>>(g);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PLIFE_OK;
}

// Applies position preset (positions) or type preset (!positions) `preset`
// to the slots of world w, with draws from `seed`; the layout must be synced.
int apply_preset(plife_physics *h, WorldHost &w, bool positions, int preset, uint64_t seed)
{
    if (preset < 0 || preset >= (positions ? (int)kPosPresets : (int)kTypePresets))
        return fail(PLIFE_ERR_ARGUMENT, "%s preset %d out of range", positions ? "position" : "type", preset);
    if (w.n == 0) return PLIFE_OK;
    if (!positions && w.max_type >= w.ntypes)
        return fail(PLIFE_ERR_STATE, "particle types reach %d but the matrix has %d types; set the particles first",
                    w.max_type, w.ntypes);
    PresetArgs g{};
    for (int s = 0; s < 6; ++s) g.base[s] = splitmix_mix(seed + (uint64_t)(kPresetStream + s) * kStream);
    for (int s = 0; s < 4; ++s) g.ubase[s] = splitmix_mix(seed + (uint64_t)(s + 1) * kStream);
    g.n = w.n;
    g.off = w.off;
    g.ntypes = w.ntypes;
    g.dim = w.st.dim;
    g.preset = preset;
    g.wrap = w.st.wrap != 0;
    const int p = h->pc, a = h->ac;
    g.px = h->px[p];
    g.py = h->py[p];
    g.pz = h->pz[p];
    g.vx = h->vx[a];
    g.vy = h->vy[a];
    g.vz = h->vz[a];
    g.type = h->type[a];
    if (positions) k_set_positions<<<ceil_div(w.n, kBlock), kBlock, 0, h->stream>>>(g);
    else k_set_types<<<ceil_div(w.n, kBlock), kBlock, 0, h->stream>>>(g);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(h->stream));
    w.nb = -1.0;             // the world's clustering is no longer known
    if (!positions) w.max_type = w.ntypes - 1;
    h->stepped = false;      // cell by-products describe the old state
    return PLIFE_OK;
}

// Overwrites k particles of world w at slots of its current order (the
// order downloads return), keeping their ids; the layout must be synced.
int set_slots(plife_physics *h, WorldHost &w, int32_t k, const int32_t *slots, const double *pos,
              const double *vel, const int32_t *types)
{
    if (k < 0) return fail(PLIFE_ERR_ARGUMENT, "k must be >= 0, got %d", k);
    if (k == 0) return PLIFE_OK;
    if (!slots || !pos || !vel || !types) return fail(PLIFE_ERR_ARGUMENT, "null array");
    std::vector<uint8_t> seen((size_t)std::max(w.n, 1), 0);
    std::vector<SlotEdit> e((size_t)k);
    const bool wrap = w.st.wrap != 0, d3 = w.st.dim == 3;
    int tmax = -1;
    for (int i = 0; i < k; ++i) {
        const int s = slots[i];
        if (s < 0 || s >= w.n) return fail(PLIFE_ERR_ARGUMENT, "slot %d outside [0, %d)", s, w.n);
        if (seen[s]) return fail(PLIFE_ERR_ARGUMENT, "slot %d given twice", s);
        seen[s] = 1;
        if (types[i] < 0 || types[i] >= w.ntypes)
            return fail(PLIFE_ERR_ARGUMENT, "edit %d has type %d outside [0, %d)", i, types[i], w.ntypes);
        e[i] = SlotEdit{s, fixed_of(pos[3 * i], wrap), fixed_of(pos[3 * i + 1], wrap),
                        d3 ? fixed_of(pos[3 * i + 2], wrap) : 0u, (float)vel[3 * i], (float)vel[3 * i + 1],
                        d3 ? (float)vel[3 * i + 2] : 0.0f, types[i]};
        tmax = std::max(tmax, types[i]);
    }
    int rc;
    if ((rc = ensure_dev(&h->d_edits, &h->edits_cap, (size_t)k, &h->retired))) return rc;
    const size_t bytes = (size_t)k * sizeof(SlotEdit);
    if ((rc = begin_uploads(h, bytes))) return rc;
    if ((rc = upload(h, h->d_edits, e.data(), bytes))) return rc;
    if ((rc = end_uploads(h))) return rc;
    const int p = h->pc, a = h->ac;
    k_set_slots<<<ceil_div(k, kBlock), kBlock, 0, h->stream>>>(h->d_edits, k, w.off, h->px[p], h->py[p],
                                                               h->pz[p], h->vx[a], h->vy[a], h->vz[a],
                                                               h->type[a]);
    LAUNCH_CHECK();
    h->stepped = false;
    if (w.max_type < w.ntypes) {
        w.max_type = std::max(w.max_type, tmax);
        return PLIFE_OK;
    }
    // a matrix shrunk below the types: the edits may be what brings them
    // back within it, so take the true maximum
    std::vector<int32_t> all((size_t)w.n);
    HIP_TRY(hipMemcpyAsync(all.data(), h->type[a] + w.off, (size_t)w.n * 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    w.max_type = *std::max_element(all.begin(), all.end());
    return PLIFE_OK;
}

// Downloads slots [off, off + n) of the current buffers.  k_pack gathers the
// requested arrays chunk by chunk into device staging, one DMA per chunk moves
// it to pinned memory, and host threads convert each chunk to the caller's
// FP64 triples as soon as its DMA has completed, while later chunks are still
// in flight.
int download_particles(plife_physics *h, int off, int n, double *pos, double *vel, int32_t *types,
                       int32_t *ids, bool d3)
{
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (n == 0) return PLIFE_OK;
    const PosBuf P = h->P(h->pc);
    const AuxBuf A = h->A(h->ac);
    PackArgs pk{};
    int na = 0;
    auto add = [&](bool want, const void *p) {
        if (!want) return -1;
        pk.src[na] = (const uint32_t *)p;
        return na++;
    };
    const int ix = add(pos, P.x), iy = add(pos, P.y), iz = add(pos && d3, P.z);
    const int ivx = add(vel, A.vx), ivy = add(vel, A.vy), ivz = add(vel && d3, A.vz);
    const int it = add(types, A.type), ii = add(ids, A.id);
    if (na == 0) return PLIFE_OK;
    const size_t words = (size_t)na * (size_t)n;
    int rc = ensure_dev(&h->dl_dev, &h->dl_cap, words);
    if (rc) return rc;
    if ((rc = ensure_pin(h, words * 4))) return rc;
    const int chunk = n >= 2 * kDlMinChunk
                          ? (int)std::min<long long>(n, ((long long)ceil_div(n, kDlChunks) + 255) / 256 * 256)
                          : n;
    const int nchunks = ceil_div(n, chunk);
    pk.na = na;
    pk.off = off;
    pk.n = n;
    pk.chunk = chunk;
    k_pack<<<ceil_div(n, kBlock), kBlock, 0, h->stream>>>(pk, h->dl_dev);
    LAUNCH_CHECK();
    uint32_t *const stage = (uint32_t *)h->pin;
    for (int c = 0; c < nchunks; ++c) {
        const size_t first = (size_t)c * chunk * na;
        const size_t len = (size_t)std::min(chunk, n - c * chunk);
        HIP_TRY(hipMemcpyAsync(stage + first, h->dl_dev + first, len * na * 4, hipMemcpyDeviceToHost,
                               h->stream));
        HIP_TRY(hipEventRecord(h->dl_ev[c], h->stream));
    }

    // chunk c, its slots [lo, hi) (relative to the chunk)
    auto convert = [&](int c, int lo, int hi) {
        const int g = c * chunk;
        const int len = std::min(chunk, n - g);
        const uint32_t *base = stage + (size_t)g * na;
        auto col = [&](int a) { return base + (size_t)a * len; };
        if (pos) {
            const uint32_t *x = col(ix), *y = col(iy), *z = d3 ? col(iz) : nullptr;
            for (int i = lo; i < hi; ++i) {
                double *o = pos + 3 * (size_t)(g + i);
                o[0] = (double)x[i] * kInvFix;
                o[1] = (double)y[i] * kInvFix;
                o[2] = z ? (double)z[i] * kInvFix : 0.0;
            }
        }
        if (vel) {
            const float *x = (const float *)col(ivx), *y = (const float *)col(ivy);
            const float *z = d3 ? (const float *)col(ivz) : nullptr;
            for (int i = lo; i < hi; ++i) {
                double *o = vel + 3 * (size_t)(g + i);
                o[0] = x[i];
                o[1] = y[i];
                o[2] = z ? (double)z[i] : 0.0;
            }
        }
        if (types) std::memcpy(types + g + lo, col(it) + lo, (size_t)(hi - lo) * 4);
        if (ids) std::memcpy(ids + g + lo, col(ii) + lo, (size_t)(hi - lo) * 4);
    };
    // This thread waits for the chunks' events in order and publishes them;
    // every thread (this one included) converts its share of each chunk.
    const int nt = n >= kDlMinChunk ? host_threads() : 1;
    std::atomic<int> ready{0};
    std::atomic<bool> failed{false};
    auto share = [&](int t, int c) {
        const int len = std::min(chunk, n - c * chunk);
        const long long lo = (long long)len * t / nt, hi = (long long)len * (t + 1) / nt;
        if (lo < hi) convert(c, (int)lo, (int)hi);
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < nt; ++t)
        ts.emplace_back([&, t]() {
            for (int c = 0; c < nchunks; ++c) {
                while (ready.load(std::memory_order_acquire) <= c) std::this_thread::yield();
                if (failed.load(std::memory_order_acquire)) return;
                share(t, c);
            }
        });
    hipError_t err = hipSuccess;
    for (int c = 0; c < nchunks; ++c) {
        err = hipEventSynchronize(h->dl_ev[c]);
        if (err != hipSuccess) {
            failed.store(true, std::memory_order_release);
            ready.store(nchunks, std::memory_order_release);
            break;
        }
        ready.store(c + 1, std::memory_order_release);
        share(0, c);
    }
    for (auto &t : ts) t.join();
    if (err != hipSuccess) return fail(PLIFE_ERR_DEVICE, "download: %s", hipGetErrorString(err));
    return PLIFE_OK;
}

int create_handle(int32_t device, plife_physics *h)
{
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(PLIFE_ERR_ARGUMENT, "device %d out of range (%d devices)", device, count);
    h->device = device;
    int rc = set_device(h);
    if (rc) return rc;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
        lds > 0)
        h->lds_limit = lds;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(PLIFE_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    for (UploadSlot &s : h->up) {
        e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (e != hipSuccess) return fail(PLIFE_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e));
    }
    for (hipEvent_t &ev : h->dl_ev) {
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return fail(PLIFE_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e));
    }
    for (hipEvent_t &ev : h->fl_ev) {
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return fail(PLIFE_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e));
    }
    for (hipEvent_t &ev : h->tune_ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) return fail(PLIFE_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e));
    }
    e = hipHostMalloc((void **)&h->stat_host, 4 * sizeof(uint64_t), hipHostMallocDefault);
    if (e != hipSuccess) return fail(PLIFE_ERR_MEMORY, "hipHostMalloc: %s", hipGetErrorString(e));
    std::memset(h->stat_host, 0, 4 * sizeof(uint64_t));
    h->stat_cap = 4;
    e = hipHostGetDevicePointer((void **)&h->stat_dev, h->stat_host, 0);
    if (e != hipSuccess) return fail(PLIFE_ERR_DEVICE, "hipHostGetDevicePointer: %s", hipGetErrorString(e));
    return PLIFE_OK;
}

void destroy_handle(plife_physics *h)
{
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_state(h);
    dev_free(h->cell_count);
    dev_free(h->cell_start);
    dev_free(h->cell_end);
    dev_free(h->hist);
    dev_free(h->bsum);
    dev_free(h->bsq);
    dev_free(h->d_pairs);
    dev_free(h->d_facc);
    dev_free(h->d_matrix);
    dev_free(h->d_worlds);
    dev_free(h->world_of);
    dev_free(h->d_src_off);
    dev_free(h->d_blocks);
    dev_free(h->d_edits);
    dev_free(h->dbg_count);
    dev_free(h->dbg_hash);
    dev_free(h->dbg_acc);
    if (h->stat_host) (void)hipHostFree(h->stat_host);
    if (h->work_host) (void)hipHostFree(h->work_host);
    for (void *p : h->retired_host) (void)hipHostFree(p);
    h->retired_host.clear();
    if (h->pin) (void)hipHostFree(h->pin);
    dev_free(h->dl_dev);
    for (hipEvent_t &ev : h->dl_ev) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    for (hipEvent_t &ev : h->fl_ev) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    for (hipEvent_t &ev : h->tune_ev) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    for (void *p : h->retired) (void)hipFree(p);
    h->retired.clear();
    for (UploadSlot &s : h->up) {
        if (s.p) (void)hipHostFree(s.p);
        if (s.done) (void)hipEventDestroy(s.done);
        s = UploadSlot{};
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

// A new world: forget the clustering measure of the old ones.
void reset_planner(plife_physics *h)
{
    std::memset(h->stat_host, 0, h->stat_cap * sizeof(uint64_t));
    h->occupancy = 0.0;
    h->clustered = false;
    for (WorldHost &w : h->worlds) w.nb = -1.0;
}

int check_single(plife_physics *h)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (h->batch) return fail(PLIFE_ERR_ARGUMENT, "batch handles take the plife_batch_* calls");
    return PLIFE_OK;
}

int check_ready(plife_physics *h)
{
    int rc = check_single(h);
    if (rc) return rc;
    if (!h->have_settings) return fail(PLIFE_ERR_STATE, "plife_set_settings has not been called");
    return PLIFE_OK;
}

int check_tuning(const plife_tuning *t)
{
    if (!t) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    if (t->list_cap < 0 || t->list_cap > 512 || !(t->stage_slack >= 0.0) ||
        t->stage_slack > 16.0 || t->direct_launch < 0 || t->direct_launch > 1 ||
        t->sparse_worlds < 0 || t->sparse_worlds > 2 || t->windowed_only < 0 ||
        t->windowed_only > 1 || t->small_step < 0 || t->small_step > 2 ||
        (t->block_homes != 0 && t->block_homes != 64 && t->block_homes != 128 &&
         t->block_homes != 256))
        return fail(PLIFE_ERR_ARGUMENT, "tuning out of range");
    return PLIFE_OK;
}

void report_tuning(const plife_physics *h, plife_tuning *t, int32_t *lds_bytes)
{
    *t = h->tune;
    // automatic fields report what the last force launch used
    if (t->list_cap == 0) t->list_cap = h->last_lcap;
    if (t->block_homes == 0) t->block_homes = h->last_bs;
    t->small_step = h->small ? 2 : 1;   // the path taken: 2 one launch, 1 multi-launch
    if (lds_bytes) *lds_bytes = (int32_t)h->last_lds;
}

// ---------------------------------------------------------------------------
// the step without allocation (plife_step_noalloc)
// ---------------------------------------------------------------------------
//
// Everything a step needs beyond the caller's state arrays lives in one
// file-scope device array, `scratch`, of fixed size: the world's constants,
// the workgroup table, the step's counters, the keys, radix buffers and
// permutation, and the cell arrays.  Each GPU the library runs on holds its
// own copy of the array until the process ends.  Host code never passes the
// array's name to a kernel: the address each device holds it at comes from
// hipGetSymbolAddress (na_scratch_base) and only that address is used.

constexpr int kNaMaxN = 262144;                  // particles per call
constexpr long long kNaMaxCells = 1LL << 20;     // cells per call (nx^dim)
constexpr int kNaMaxBlocks = kNaMaxN / 64;       // force workgroups (64 homes or more each)

struct NaLayout {
    size_t world, counters, blocks, work, keys0, kA, kB, vA, vB, skeys, count, start, end, hist,
        bsum, bsq, total;
};

constexpr size_t na_align(size_t o) { return (o + 255) & ~(size_t)255; }

constexpr NaLayout na_layout()
{
    NaLayout L{};
    size_t o = 0;
    L.world = o;    o = na_align(o + sizeof(WorldDev));
    L.counters = o; o = na_align(o + 8 * sizeof(unsigned long long));
    L.blocks = o;   o = na_align(o + (size_t)kNaMaxBlocks * sizeof(uint2));
    L.work = o;     o = na_align(o + ((size_t)kNaMaxBlocks + 1) * sizeof(uint32_t));
    L.keys0 = o;    o = na_align(o + (size_t)kNaMaxN * 4);
    L.kA = o;       o = na_align(o + (size_t)kNaMaxN * 4);
    L.kB = o;       o = na_align(o + (size_t)kNaMaxN * 4);
    L.vA = o;       o = na_align(o + (size_t)kNaMaxN * 4);
    L.vB = o;       o = na_align(o + (size_t)kNaMaxN * 4);
    L.skeys = o;    o = na_align(o + (size_t)kNaMaxN * 4);
    L.count = o;    o = na_align(o + (size_t)kNaMaxCells * 4);
    L.start = o;    o = na_align(o + (size_t)kNaMaxCells * 4);
    L.end = o;      o = na_align(o + (size_t)kNaMaxCells * 4);
    // radix digit histograms: at most 2^8 digits per tile (passes of <= 8 bits)
    L.hist = o;     o = na_align(o + (size_t)256 * ((kNaMaxN + kRadixTile - 1) / kRadixTile) * 4);
    L.bsum = o;     o = na_align(o + (size_t)4096 * 4);
    L.bsq = o;      o = na_align(o + (size_t)((kNaMaxCells + kScanTile - 1) / kScanTile) * 8);
    L.total = o;
    return L;
}

constexpr NaLayout kNa = na_layout();

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNa.total];

namespace {

std::mutex g_na_mu;
unsigned char *g_na_base[64] = {};

// The address device `dev` holds `scratch` at, looked up once per device
// with that device current.
int na_scratch_base(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= 64) return fail(PLIFE_ERR_ARGUMENT, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_base[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PLIFE_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_na_base[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_na_base[dev];
    return PLIFE_OK;
}

// Per call: the world's constants, the workgroup table (world 0, first home
// b * bs), and zeros where the step expects them -- the counters, the tag
// slot of the workgroup work record, and the first ncells cell counts (a
// previous call may have stopped between binning and the cell scan).
__global__ void k_na_setup(WorldDev wd, WorldDev *world, uint2 *blocks, int nblocks, int bs,
                           unsigned long long *counters, uint32_t *work, uint32_t *cell_count,
                           int ncells)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *world = wd;
    if (i < 8) counters[i] = 0ull;
    if (i < nblocks) blocks[i] = make_uint2(0u, (uint32_t)(i * bs));
    if (i == 0) work[nblocks] = 0u;
    for (int c = i; c < ncells; c += gridDim.x * blockDim.x) cell_count[c] = 0u;
}

__global__ void k_na_copy_pairs(const unsigned long long *counters, unsigned long long *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = counters[0];
}

int noalloc_step(const plife_settings *s, int32_t ntypes, const double *matrix, int32_t n,
                 const uint32_t *px, const uint32_t *py, const uint32_t *pz, const float *vx,
                 const float *vy, const float *vz, const int32_t *type, const int32_t *id,
                 uint32_t *opx, uint32_t *opy, uint32_t *opz, float *ovx, float *ovy, float *ovz,
                 int32_t *otype, int32_t *oid, uint64_t *pairs, hipStream_t stream)
{
    if (!s || !matrix) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    if (s->dim != 2 && s->dim != 3) return fail(PLIFE_ERR_ARGUMENT, "dim must be 2 or 3, got %d", s->dim);
    if (!(s->rmax > 0.0) || !(s->rmax <= 1.0))
        return fail(PLIFE_ERR_ARGUMENT, "rmax must be in (0, 1], got %g", s->rmax);
    if (!(s->beta > 0.0) || !(s->beta < 1.0))
        return fail(PLIFE_ERR_ARGUMENT, "beta must be in (0, 1), got %g", s->beta);
    if (!std::isfinite(s->dt) || !std::isfinite(s->force) || !std::isfinite(s->friction))
        return fail(PLIFE_ERR_ARGUMENT, "dt, force and friction must be finite");
    if (ntypes < 1 || ntypes > kMaxTypes)
        return fail(PLIFE_ERR_ARGUMENT, "matrix size must be in [1, %d], got %d", kMaxTypes, ntypes);
    if (n < 0 || n > kNaMaxN)
        return fail(PLIFE_ERR_ARGUMENT, "n = %d exceeds the scratch limit of %d particles", n, kNaMaxN);
    const Grid gr = grid_of(*s);
    if (gr.ncells > kNaMaxCells)
        return fail(PLIFE_ERR_ARGUMENT, "%lld cells (rmax %g, dim %d) exceed the scratch limit of %lld",
                    gr.ncells, s->rmax, s->dim, kNaMaxCells);
    const bool d3 = s->dim == 3;
    if (n > 0 && (!px || !py || !vx || !vy || !type || !id || !opx || !opy || !ovx || !ovy ||
                  !otype || !oid || (d3 && (!pz || !vz || !opz || !ovz))))
        return fail(PLIFE_ERR_ARGUMENT, "null state array");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    unsigned char *base = nullptr;
    int rc = na_scratch_base(dev, &base);
    if (rc) return rc;

    WorldHost wh;
    wh.st = *s;
    wh.ntypes = ntypes;
    wh.n = n;
    wh.off = 0;
    const WorldDev wd = world_dev(wh, 0u, 0);
    const int nc = (int)gr.ncells;
    const double avg = (double)n / (double)nc;
    const ForcePlan fp = plan_force_for(s->dim, avg, avg + 1.0, false, plife_tuning{}, false, ntypes);
    const int nblocks = ceil_div(n, fp.bs);
    const RadixPlan plan = radix_plan(nc, n);

    // a transient view of the scratch as the step helpers' buffers (no
    // allocation: plain pointers into the array)
    plife_physics t;
    t.stream = stream;
    t.dim = s->dim;
    t.n = n;
    t.ncells = nc;
    t.ntypes_max = ntypes;
    t.d_worlds = reinterpret_cast<WorldDev *>(base + kNa.world);
    t.d_pairs = reinterpret_cast<unsigned long long *>(base + kNa.counters);
    t.d_blocks = reinterpret_cast<uint2 *>(base + kNa.blocks);
    t.keys0 = reinterpret_cast<uint32_t *>(base + kNa.keys0);
    t.kA = reinterpret_cast<uint32_t *>(base + kNa.kA);
    t.kB = reinterpret_cast<uint32_t *>(base + kNa.kB);
    t.vA = reinterpret_cast<uint32_t *>(base + kNa.vA);
    t.vB = reinterpret_cast<uint32_t *>(base + kNa.vB);
    t.skeys = reinterpret_cast<uint32_t *>(base + kNa.skeys);
    t.cell_count = reinterpret_cast<uint32_t *>(base + kNa.count);
    t.cell_start = reinterpret_cast<uint32_t *>(base + kNa.start);
    t.cell_end = reinterpret_cast<uint32_t *>(base + kNa.end);
    t.hist = reinterpret_cast<uint32_t *>(base + kNa.hist);
    t.bsum = reinterpret_cast<uint32_t *>(base + kNa.bsum);
    t.bsum_cap = (kNa.bsq - kNa.bsum) / 4;
    t.bsq = reinterpret_cast<unsigned long long *>(base + kNa.bsq);
    t.bsq_cap = (kNa.total - kNa.bsq) / 8;
    uint32_t *work = reinterpret_cast<uint32_t *>(base + kNa.work);
    if ((size_t)ceil_div(nc, kScanTile) + (size_t)ceil_div((long long)plan.hist_len, kScanTile) >
            t.bsum_cap ||
        plan.hist_len > (kNa.bsum - kNa.hist) / 4 || nblocks > kNaMaxBlocks)
        return fail(PLIFE_ERR_STATE, "scratch layout too small for this world");

    const int setup_threads = std::max(nblocks + 1, 8);
    k_na_setup<<<std::min(ceil_div(std::max(setup_threads, nc), kBlock), 1024), kBlock, 0, stream>>>(
        wd, t.d_worlds, t.d_blocks, nblocks, fp.bs, t.d_pairs, work, t.cell_count, nc);
    LAUNCH_CHECK();
    if (n > 0) {
        // 1. cell keys, histogram and first digit histogram; counters [0, 3)
        //    cleared, tag in [3] (Physics.java:319-340)
#define PLIFE_NA_BIN(DD)                                                                          \
    do {                                                                                          \
        if (d3) k_bin_tile<3, DD><<<plan.ntiles, 1024, 0, stream>>>(px, py, pz, n, t.d_worlds,   \
                nullptr, plan.ntiles, t.keys0, t.cell_count, t.hist, t.d_pairs, 3, 1ull, nullptr); \
        else k_bin_tile<2, DD><<<plan.ntiles, 1024, 0, stream>>>(px, py, pz, n, t.d_worlds,      \
                nullptr, plan.ntiles, t.keys0, t.cell_count, t.hist, t.d_pairs, 3, 1ull, nullptr); \
    } while (0)
        switch (plan.D) {
            case 1: PLIFE_NA_BIN(1); break;
            case 2: PLIFE_NA_BIN(2); break;
            case 3: PLIFE_NA_BIN(3); break;
            case 4: PLIFE_NA_BIN(4); break;
            case 5: PLIFE_NA_BIN(5); break;
            case 6: PLIFE_NA_BIN(6); break;
            case 7: PLIFE_NA_BIN(7); break;
            default: PLIFE_NA_BIN(8); break;
        }
#undef PLIFE_NA_BIN
        LAUNCH_CHECK();
        // cell scan (clears the counts) and first digit scan
        if ((rc = scan_jobs(&t, ScanJob{t.cell_count, nc, t.cell_start, t.cell_end, 1, nullptr,
                                        t.d_pairs + 1},
                            ScanJob{t.hist, (int)plan.hist_len, t.hist, nullptr, 0, nullptr, nullptr})))
            return rc;
        // 2. the stable sort: the reference's counting-sort permutation
        if ((rc = radix_sort(&t, n, plan))) return rc;
        // 3. force + integrate into the caller's output arrays
        ForceArgs a{};
        a.perm = t.perm;
        a.sx = px;
        a.sy = py;
        a.sz = pz;
        a.vx = vx;
        a.vy = vy;
        a.vz = vz;
        a.type = type;
        a.id = id;
        a.ox = opx;
        a.oy = opy;
        a.oz = opz;
        a.ovx = ovx;
        a.ovy = ovy;
        a.ovz = ovz;
        a.otype = otype;
        a.oid = oid;
        a.skeys = t.skeys;
        a.cell_start = t.cell_start;
        a.cell_end = t.cell_end;
        a.matrix = matrix;
        a.worlds = t.d_worlds;
        a.blocks = t.d_blocks;
        a.ntypes_max = ntypes;
        a.m_in_lds = ntypes <= kMaxTypesLds ? 1 : 0;
        a.cap = fp.cap;
        a.lcap = fp.lcap;
        a.max_cells = fp.max_cells;
        a.wpool = fp.wpool;
        a.pair_total = t.d_pairs;
        a.pair_world = t.d_pairs + 2;
        a.block_work = work;
        a.work_tag = 1u;
        a.nb_total = nblocks;
        t.seg[s->wrap ? 0 : 2] = nblocks;
        if ((rc = launch_force_any(&t, a, fp.bs, fp.lds))) return rc;
    }
    if (pairs) {
        k_na_copy_pairs<<<1, 64, 0, stream>>>(t.d_pairs, reinterpret_cast<unsigned long long *>(pairs));
        LAUNCH_CHECK();
    }
    return PLIFE_OK;
}

WorldHost *find_world(plife_physics *h, int32_t world_id)
{
    auto it = h->index_of.find(world_id);
    if (it == h->index_of.end()) return nullptr;
    WorldHost &w = h->worlds[it->second];
    return w.live ? &w : nullptr;
}

}  // namespace

extern "C" {

const char *plife_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// single world
// ---------------------------------------------------------------------------

int plife_create(int32_t device, plife_physics **out)
{
    if (!out) return fail(PLIFE_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    plife_physics *h = new plife_physics();
    int rc = create_handle(device, h);
    if (rc) {
        destroy_handle(h);
        delete h;
        return rc;
    }
    h->worlds.emplace_back();
    h->index_of[0] = 0;
    *out = h;
    g_last_error.clear();
    return PLIFE_OK;
}

int plife_destroy(plife_physics *h)
{
    if (!h) return PLIFE_OK;
    if (h->batch) return fail(PLIFE_ERR_ARGUMENT, "batch handles take plife_batch_destroy");
    destroy_handle(h);
    delete h;
    return PLIFE_OK;
}

int plife_set_settings(plife_physics *h, const plife_settings *s, int32_t ntypes,
                       const double *matrix)
{
    int rc = check_single(h);
    if (rc) return rc;
    if ((rc = check_settings(s, ntypes, matrix))) return rc;
    store_settings(h, h->worlds[0], s, ntypes, matrix);
    h->dim = s->dim;
    h->have_settings = true;
    return PLIFE_OK;
}

int plife_set_particles(plife_physics *h, int32_t n, const double *pos, const double *vel,
                        const int32_t *types)
{
    int rc = check_ready(h);
    if (rc) return rc;
    WorldHost &w = h->worlds[0];
    if ((rc = check_particles(n, pos, vel, types, w.ntypes))) return rc;
    if ((rc = set_device(h))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if ((rc = ensure_capacity(h, n, false))) return rc;
    if ((rc = upload_particles(h, 0, n, pos, vel, types, w.st.wrap != 0, w.st.dim == 3))) return rc;
    w.n = n;
    w.off = 0;
    w.max_type = n > 0 ? *std::max_element(types, types + n) : -1;
    h->used = n;
    h->layout_dirty = true;
    h->stepped = false;
    reset_planner(h);
    return PLIFE_OK;
}

int plife_generate_particles(plife_physics *h, int32_t n, uint64_t seed)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (n < 0) return fail(PLIFE_ERR_ARGUMENT, "n must be >= 0, got %d", n);
    WorldHost &w = h->worlds[0];
    if ((rc = set_device(h))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if ((rc = ensure_capacity(h, n, false))) return rc;
    if ((rc = generate_particles(h, 0, n, seed, w.ntypes, w.st.dim))) return rc;
    w.n = n;
    w.off = 0;
    w.max_type = n > 0 ? w.ntypes - 1 : -1;
    h->used = n;
    h->layout_dirty = true;
    h->stepped = false;
    reset_planner(h);
    return PLIFE_OK;
}

int plife_set_positions(plife_physics *h, int32_t preset, uint64_t seed)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = set_device(h))) return rc;
    if ((rc = sync_layout(h))) return rc;
    if ((rc = apply_preset(h, h->worlds[0], true, preset, seed))) return rc;
    reset_planner(h);
    return PLIFE_OK;
}

int plife_set_particle_slots(plife_physics *h, int32_t k, const int32_t *slots, const double *pos,
                             const double *vel, const int32_t *types)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = set_device(h))) return rc;
    if ((rc = sync_layout(h))) return rc;
    return set_slots(h, h->worlds[0], k, slots, pos, vel, types);
}

int plife_set_types(plife_physics *h, int32_t preset, uint64_t seed)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if ((rc = set_device(h))) return rc;
    if ((rc = sync_layout(h))) return rc;
    return apply_preset(h, h->worlds[0], false, preset, seed);
}

int plife_generate_matrix(int32_t preset, int32_t size, uint64_t seed, double *out)
{
    if (!out) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    if (size < 1 || size > kMaxTypes)
        return fail(PLIFE_ERR_ARGUMENT, "size must be in [1, %d], got %d", kMaxTypes, size);
    if (preset < 0 || preset > 6) return fail(PLIFE_ERR_ARGUMENT, "matrix preset %d out of range", preset);
    const int n = size;
    auto at = [&](int i, int j) -> double & { return out[(size_t)i * n + j]; };
    const bool random = preset == 0 || preset == 1;
    // DefaultMatrix.randomize (DefaultMatrix.java:24-30): 2u - 1 for every entry in row order
    const uint64_t base = splitmix_mix(seed + 5 * kStream);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const uint64_t k = (uint64_t)i * n + j;
            at(i, j) = random ? 2.0 * ((double)(splitmix_mix(base + (k + 1) * kGamma) >> 40) * 0x1p-24) - 1.0
                              : 0.0;
        }
    switch (preset) {
        case 1:     // symmetry: the upper triangle copies the lower one
            for (int i = 0; i < n; ++i)
                for (int j = i; j < n; ++j) at(i, j) = at(j, i);
            break;
        case 2:     // chains: 1 on and beside the diagonal (cyclically), -1 elsewhere
        case 3:     // chains 2: 1 on the diagonal, 0.2 beside it, -1 elsewhere
        case 4:     // chains 3: as chains 2 with 0 elsewhere
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    const bool next = j == (i + 1) % n || j == (i + n - 1) % n;
                    if (preset == 2) at(i, j) = (j == i || next) ? 1.0 : -1.0;
                    else at(i, j) = j == i ? 1.0 : (next ? 0.2 : (preset == 3 ? -1.0 : 0.0));
                }
            break;
        case 5:     // snakes: 1 on the diagonal, then 0.2 on the next column (cyclically)
            for (int i = 0; i < n; ++i) {
                at(i, i) = 1.0;
                at(i, (i + 1) % n) = 0.2;
            }
            break;
        default:    // random, zero
            break;
    }
    return PLIFE_OK;
}

int plife_get_particles(plife_physics *h, int32_t n, double *pos, double *vel, int32_t *types,
                        int32_t *ids)
{
    int rc = check_single(h);
    if (rc) return rc;
    const WorldHost &w = h->worlds[0];
    if (n != w.n) return fail(PLIFE_ERR_ARGUMENT, "n = %d but the handle holds %d particles", n, w.n);
    if ((rc = set_device(h))) return rc;
    return download_particles(h, w.off, n, pos, vel, types, ids, h->have_settings && h->dim == 3);
}

int plife_get_count(plife_physics *h, int32_t *n)
{
    int rc = check_single(h);
    if (rc) return rc;
    if (!n) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    *n = h->worlds[0].n;
    return PLIFE_OK;
}

int plife_get_steps(plife_physics *h, int64_t *steps)
{
    int rc = check_single(h);
    if (rc) return rc;
    if (!steps) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    *steps = h->worlds[0].steps;
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
    const Grid g = grid_of(h->worlds[0].st);
    if (nx) *nx = g.nx;
    if (ncells) *ncells = (int32_t)g.ncells;
    return PLIFE_OK;
}

int plife_get_cells(plife_physics *h, int32_t *keys, int32_t *cell_end, int32_t *perm,
                    int32_t ncells)
{
    int rc = check_single(h);
    if (rc) return rc;
    if (!h->stepped) return fail(PLIFE_ERR_STATE, "no step has run since the state was set");
    if (ncells != h->last_ncells)
        return fail(PLIFE_ERR_ARGUMENT, "ncells = %d but the last step used %lld cells", ncells, h->last_ncells);
    if ((rc = set_device(h))) return rc;
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
    return pair_total(h, pairs);
}

int plife_set_tuning(plife_physics *h, const plife_tuning *t)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    int rc = check_tuning(t);
    if (rc) return rc;
    h->tune = *t;
    h->small_meas = ~0ull;             // re-time the paths and shapes under the new tuning
    h->bs_meas = ~0ull;
    h->auto_bs = 0;
    h->bs_seen.clear();
    return PLIFE_OK;
}

int plife_get_tuning(plife_physics *h, plife_tuning *t, int32_t *lds_bytes)
{
    if (!h || !t) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    report_tuning(h, t, lds_bytes);
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

int plife_step_noalloc(const plife_settings *s, int32_t ntypes, const double *matrix, int32_t n,
                       const uint32_t *px, const uint32_t *py, const uint32_t *pz, const float *vx,
                       const float *vy, const float *vz, const int32_t *type, const int32_t *id,
                       uint32_t *opx, uint32_t *opy, uint32_t *opz, float *ovx, float *ovy,
                       float *ovz, int32_t *otype, int32_t *oid, uint64_t *pairs, void *stream)
{
    return noalloc_step(s, ntypes, matrix, n, px, py, pz, vx, vy, vz, type, id, opx, opy, opz, ovx,
                        ovy, ovz, otype, oid, pairs, static_cast<hipStream_t>(stream));
}

int plife_get_accelerations(plife_physics *h, int32_t n, float *acc)
{
    if (!h) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!acc) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    if (!h->debug || !h->stepped) return fail(PLIFE_ERR_STATE, "debug mode was off or no step has run");
    if (n != h->n) return fail(PLIFE_ERR_ARGUMENT, "n = %d but the handle holds %d particles", n, h->n);
    int rc = set_device(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (n == 0) return PLIFE_OK;
    HIP_TRY(hipMemcpy(acc, h->dbg_acc, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return PLIFE_OK;
}

int plife_time_steps(plife_physics *h, int32_t nsteps, double *ms)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!ms || nsteps < 0) return fail(PLIFE_ERR_ARGUMENT, "bad argument");
    if ((rc = set_device(h))) return rc;
    return time_steps(h, nsteps, ms);
}

int plife_profile(plife_physics *h, int32_t nsteps, plife_timing *out)
{
    int rc = check_ready(h);
    if (rc) return rc;
    if (!out || nsteps <= 0) return fail(PLIFE_ERR_ARGUMENT, "bad argument");
    if ((rc = set_device(h))) return rc;
    return profile(h, nsteps, out);
}

// ---------------------------------------------------------------------------
// batches of worlds
// ---------------------------------------------------------------------------

int plife_batch_create(int32_t device, plife_batch **out)
{
    if (!out) return fail(PLIFE_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    plife_batch *b = new plife_batch();
    b->batch = true;
    int rc = create_handle(device, b);
    if (rc) {
        destroy_handle(b);
        delete b;
        return rc;
    }
    *out = b;
    g_last_error.clear();
    return PLIFE_OK;
}

int plife_batch_destroy(plife_batch *b)
{
    if (!b) return PLIFE_OK;
    destroy_handle(b);
    delete b;
    return PLIFE_OK;
}

int plife_batch_add_world(plife_batch *b, const plife_settings *s, int32_t ntypes,
                          const double *matrix, int32_t n, const double *pos, const double *vel,
                          const int32_t *types, int32_t *world_id)
{
    if (!b || !world_id) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    int rc = check_settings(s, ntypes, matrix);
    if (rc) return rc;
    if ((rc = check_particles(n, pos, vel, types, ntypes))) return rc;
    int live = 0;
    for (const WorldHost &w : b->worlds) live += w.live ? 1 : 0;
    if (live > 0 && s->dim != b->dim)
        return fail(PLIFE_ERR_ARGUMENT, "the batch holds %dD worlds, got dim %d", b->dim, s->dim);
    if ((int)b->worlds.size() >= kMaxWorlds)
        return fail(PLIFE_ERR_ARGUMENT, "a batch holds at most %d worlds", kMaxWorlds);
    if ((long long)b->used + n > 0x7FFFFFFFLL)
        return fail(PLIFE_ERR_ARGUMENT, "a batch holds at most 2^31 - 1 particles");
    if ((rc = set_device(b))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if ((rc = ensure_capacity(b, b->used + n, true))) return rc;
    if ((rc = upload_particles(b, b->used, n, pos, vel, types, s->wrap != 0, s->dim == 3))) return rc;
    WorldHost w;
    w.max_type = n > 0 ? *std::max_element(types, types + n) : -1;
    w.id = b->next_id++;
    w.st = *s;
    w.ntypes = ntypes;
    w.matrix.assign(matrix, matrix + (size_t)ntypes * ntypes);
    w.n = n;
    w.off = b->used;
    b->used += n;
    b->dim = s->dim;
    b->index_of[w.id] = (int)b->worlds.size();
    b->worlds.push_back(std::move(w));
    b->layout_dirty = true;
    *world_id = b->worlds.back().id;
    return PLIFE_OK;
}

int plife_batch_set_world_positions(plife_batch *b, int32_t world_id, int32_t preset, uint64_t seed)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!find_world(b, world_id)) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    int rc = set_device(b);
    if (rc) return rc;
    if ((rc = sync_layout(b))) return rc;
    return apply_preset(b, *find_world(b, world_id), true, preset, seed);
}

int plife_batch_set_world_slots(plife_batch *b, int32_t world_id, int32_t k, const int32_t *slots,
                                const double *pos, const double *vel, const int32_t *types)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!find_world(b, world_id)) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    int rc = set_device(b);
    if (rc) return rc;
    if ((rc = sync_layout(b))) return rc;
    return set_slots(b, *find_world(b, world_id), k, slots, pos, vel, types);
}

int plife_batch_set_world_types(plife_batch *b, int32_t world_id, int32_t preset, uint64_t seed)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (!find_world(b, world_id)) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    int rc = set_device(b);
    if (rc) return rc;
    if ((rc = sync_layout(b))) return rc;
    return apply_preset(b, *find_world(b, world_id), false, preset, seed);
}

int plife_batch_add_generated_world(plife_batch *b, const plife_settings *s, int32_t ntypes,
                                    const double *matrix, int32_t n, uint64_t seed,
                                    int32_t *world_id)
{
    if (!b || !world_id) return fail(PLIFE_ERR_ARGUMENT, "null argument");
    int rc = check_settings(s, ntypes, matrix);
    if (rc) return rc;
    if (n < 0) return fail(PLIFE_ERR_ARGUMENT, "n must be >= 0, got %d", n);
    int live = 0;
    for (const WorldHost &w : b->worlds) live += w.live ? 1 : 0;
    if (live > 0 && s->dim != b->dim)
        return fail(PLIFE_ERR_ARGUMENT, "the batch holds %dD worlds, got dim %d", b->dim, s->dim);
    if ((int)b->worlds.size() >= kMaxWorlds)
        return fail(PLIFE_ERR_ARGUMENT, "a batch holds at most %d worlds", kMaxWorlds);
    if ((long long)b->used + n > 0x7FFFFFFFLL)
        return fail(PLIFE_ERR_ARGUMENT, "a batch holds at most 2^31 - 1 particles");
    if ((rc = set_device(b))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if ((rc = ensure_capacity(b, b->used + n, true))) return rc;
    if ((rc = generate_particles(b, b->used, n, seed, ntypes, s->dim))) return rc;
    WorldHost w;
    w.max_type = n > 0 ? ntypes - 1 : -1;
    w.id = b->next_id++;
    w.st = *s;
    w.ntypes = ntypes;
    w.matrix.assign(matrix, matrix + (size_t)ntypes * ntypes);
    w.n = n;
    w.off = b->used;
    b->used += n;
    b->dim = s->dim;
    b->index_of[w.id] = (int)b->worlds.size();
    b->worlds.push_back(std::move(w));
    b->layout_dirty = true;
    *world_id = b->worlds.back().id;
    return PLIFE_OK;
}

int plife_batch_remove_world(plife_batch *b, int32_t world_id)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    WorldHost *w = find_world(b, world_id);
    if (!w) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    w->live = false;
    b->layout_dirty = true;
    return PLIFE_OK;
}

int plife_batch_set_world_settings(plife_batch *b, int32_t world_id, const plife_settings *s,
                                   int32_t ntypes, const double *matrix)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    int rc = check_settings(s, ntypes, matrix);
    if (rc) return rc;
    WorldHost *w = find_world(b, world_id);
    if (!w) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    if (s->dim != b->dim) return fail(PLIFE_ERR_ARGUMENT, "the batch holds %dD worlds", b->dim);
    if (s->wrap != w->st.wrap || ntypes < w->ntypes) {
        // types already on the device must stay in range; wrap decides the workgroup order
        if (ntypes < w->ntypes) return fail(PLIFE_ERR_ARGUMENT, "the matrix may not shrink below the world's types");
        b->layout_dirty = true;
    }
    store_settings(b, *w, s, ntypes, matrix);
    return PLIFE_OK;
}

int plife_batch_update(plife_batch *b, int32_t nsteps)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    if (nsteps < 0) return fail(PLIFE_ERR_ARGUMENT, "nsteps must be >= 0");
    int rc = set_device(b);
    if (rc) return rc;
    b->stop.store(0);
    return run_steps(b, nsteps, true);
}

int plife_batch_force_update_stop(plife_batch *b) { return plife_force_update_stop(b); }

int plife_batch_synchronize(plife_batch *b) { return plife_synchronize(b); }

int plife_batch_get_world(plife_batch *b, int32_t world_id, int32_t n, double *pos, double *vel,
                          int32_t *types, int32_t *ids)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    const WorldHost *w = find_world(b, world_id);
    if (!w) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    if (n != w->n) return fail(PLIFE_ERR_ARGUMENT, "n = %d but world %d holds %d particles", n, world_id, w->n);
    int rc = set_device(b);
    if (rc) return rc;
    return download_particles(b, w->off, n, pos, vel, types, ids, b->dim == 3);
}

int plife_batch_world_info(plife_batch *b, int32_t world_id, int32_t *n, int64_t *pairs,
                           int64_t *steps)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    const WorldHost *w = find_world(b, world_id);
    if (!w) return fail(PLIFE_ERR_ARGUMENT, "no world %d in the batch", world_id);
    if (n) *n = w->n;
    if (steps) *steps = w->steps;
    if (pairs) {
        *pairs = 0;
        const int idx = b->index_of[world_id];
        // counters are indexed by the layout of the last step
        if (w->steps > 0 && !b->layout_dirty && b->d_pairs && 2 + idx < (int)b->pairs_cap) {
            int rc = set_device(b);
            if (rc) return rc;
            unsigned long long v = 0;
            HIP_TRY(hipMemcpyAsync(&v, b->d_pairs + 2 + idx, 8, hipMemcpyDeviceToHost, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
            *pairs = (int64_t)v;
        }
    }
    return PLIFE_OK;
}

int plife_batch_count(plife_batch *b, int32_t *nworlds, int64_t *nparticles)
{
    if (!b) return fail(PLIFE_ERR_ARGUMENT, "null handle");
    int32_t w = 0;
    int64_t p = 0;
    for (const WorldHost &wh : b->worlds)
        if (wh.live) {
            ++w;
            p += wh.n;
        }
    if (nworlds) *nworlds = w;
    if (nparticles) *nparticles = p;
    return PLIFE_OK;
}

int plife_batch_set_tuning(plife_batch *b, const plife_tuning *t) { return plife_set_tuning(b, t); }

int plife_batch_get_tuning(plife_batch *b, plife_tuning *t, int32_t *lds_bytes)
{
    return plife_get_tuning(b, t, lds_bytes);
}

int plife_batch_time_steps(plife_batch *b, int32_t nsteps, double *ms)
{
    if (!b || !ms || nsteps < 0) return fail(PLIFE_ERR_ARGUMENT, "bad argument");
    int rc = set_device(b);
    if (rc) return rc;
    return time_steps(b, nsteps, ms);
}

int plife_batch_profile(plife_batch *b, int32_t nsteps, plife_timing *out)
{
    if (!b || !out || nsteps <= 0) return fail(PLIFE_ERR_ARGUMENT, "bad argument");
    int rc = set_device(b);
    if (rc) return rc;
    return profile(b, nsteps, out);
}

}  // extern "C"
