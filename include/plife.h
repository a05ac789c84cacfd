/*
 * plife.h -- C ABI of the MI355X (gfx950) particle-life physics library.
 *
 * This is the drop-in boundary for the reference's per-frame physics step,
 * Physics.update() (src/main/java/com/particle_life/backend/Physics.java:112-134).
 * A Java GpuPhysics (JNI) or any other host binds these entry points; plain
 * pointers and sizes only.  Host arrays are caller-owned and only read/written
 * during the call; device buffers are owned by the handle.
 *
 * Conventions
 *   - Every entry point returns an int status: PLIFE_OK (0) on success, a
 *     negative PLIFE_ERR_* code on failure (plife_last_error() then returns a
 *     thread-local message).  Nothing throws across the ABI.
 *   - A handle is not thread-safe: all calls come from one thread (the app's
 *     physics Loop thread), except plife_force_update_stop().
 *   - Coordinates are in the reference's unit box: [0,1) periodic (wrap) or
 *     [0,1] clamped.  Positions and velocities cross the boundary as FP64
 *     (x,y,z) triples, the layout of app/PhysicsSnapshot.java:11-13; the
 *     device holds positions as 32-bit fixed point (x = k * 2^-32) and
 *     velocities as FP32.
 */
#ifndef PLIFE_H
#define PLIFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLIFE_OK 0
#define PLIFE_STOPPED 1          /* plife_update returned early after a stop request */
#define PLIFE_ERR_ARGUMENT (-1)  /* Java: IllegalArgumentException */
#define PLIFE_ERR_STATE (-2)     /* call order (e.g. update before settings) */
#define PLIFE_ERR_DEVICE (-3)    /* HIP runtime failure; Java: RuntimeException */
#define PLIFE_ERR_MEMORY (-4)    /* device allocation failed */

typedef struct plife_physics plife_physics;

/* PhysicsSettings (backend/PhysicsSettings.java:8-38) plus the accelerator's
 * beta (app/Main.java:276) and the world dimension.  dim 2 is the reference;
 * dim 3 is an extension (27-cell stencil). */
typedef struct {
    int32_t dim;       /* 2 or 3 */
    int32_t wrap;      /* 1 periodic [0,1), 0 clamped [0,1] */
    double rmax;       /* interaction radius, 0 < rmax <= 1; cell size = rmax */
    double friction;   /* per-1/60 s coefficient; factor = friction^(60*dt) */
    double force;      /* force scale */
    double dt;         /* time step (s) */
    double beta;       /* repulsion zone, 0 < beta < 1 */
} plife_settings;

/* Per-phase device timings from plife_profile (milliseconds per step). */
typedef struct {
    double step_ms;     /* wall time per step, event-to-event over all steps */
    double bin_ms;      /* cell keys + cell histogram + cell scan */
    double sort_ms;     /* stable radix sort passes */
    double gather_ms;   /* 0: the force kernel reads the pre-sort state through the permutation */
    double force_ms;    /* force + integrate kernel (average launch duration) */
    int64_t pairs;      /* interacting ordered pairs in the last step */
    int32_t steps;
} plife_timing;

/* Replaces the device-side half of `new Physics(...)` (Physics.java:65-80)
 * and Physics.kill() (Physics.java:156-158). */
int plife_create(int32_t device, plife_physics **out);
int plife_destroy(plife_physics *h);

/* Reads PhysicsSettings + Matrix (backend/Matrix.java:3-8, DefaultMatrix.java:39-41,
 * row = type of the acted-on particle) + beta.  Cheap to call every frame
 * (app/Main.java:292-295 overwrites dt per frame); the matrix is re-uploaded
 * only when it changed.  ntypes in [1, 256] (the app caps it at 256,
 * app/Main.java:745).  The matrix may shrink below the particles' types, as
 * Physics.setMatrixSize does before its ensureTypes (Physics.java:238-272);
 * until particles within the new size are set, plife_update and
 * plife_set_types fail with PLIFE_ERR_STATE. */
int plife_set_settings(plife_physics *h, const plife_settings *s, int32_t ntypes,
                       const double *matrix);

/* Uploads Particle[] (backend/Particle.java:5-9): pos/vel are [3n] (x,y,z),
 * types [n] in [0, ntypes).  Positions are first brought into range as
 * Physics.ensurePosition does (Physics.java:499-505).  Particle ids are the
 * array indices of this call and travel with the particles. */
int plife_set_particles(plife_physics *h, int32_t n, const double *pos, const double *vel,
                        const int32_t *types);

/* Downloads the state in the current device array order (cell-sorted after
 * a step, like Physics.particles after makeContainers' swap), for the
 * snapshot (app/PhysicsSnapshot.java:37-61) or Physics.particles.  Any output
 * pointer may be NULL; only the arrays asked for cross PCIe.  Synchronises the
 * handle's stream. */
int plife_get_particles(plife_physics *h, int32_t n, double *pos, double *vel, int32_t *types,
                        int32_t *ids);

/* In-place edits between steps -- the app's cursor tools change a few
 * particles (app/Main.java:543-580) -- without re-uploading the rest: the
 * particle at slots[i] of the current order (the order plife_get_particles
 * returns) gets pos[3i..3i+2], vel[3i..3i+2] and types[i], wrapped or
 * clamped as plife_set_particles does; ids are kept.  Slots must be distinct
 * and in [0, n), types in [0, ntypes). */
int plife_set_particle_slots(plife_physics *h, int32_t k, const int32_t *slots, const double *pos,
                             const double *vel, const int32_t *types);

int plife_get_count(plife_physics *h, int32_t *n);

/* Seeded device-side world generation: n particles of the app's default
 * presets -- uniform positions (backend/DefaultPositionSetter.java), uniform
 * types in [0, ntypes) (backend/DefaultTypeSetter.java), velocity 0
 * (Physics.java:297-303) -- drawn on the GPU from counter-based SplitMix64
 * streams of `seed` (positions on the 2^-24 grid; ids 0..n-1), replacing the
 * particles as plife_set_particles does, with no host transfer.  The
 * reference draws from unseeded Math.random; this is its reproducible form. */
int plife_generate_particles(plife_physics *h, int32_t n, uint64_t seed);

/* Physics.setPositions (Physics.java:166-168) with one of the app's position
 * presets (app/PositionSetterProvider.java:17-119), numbered in the app's
 * order: 0 centered, 1 uniform, 2 uniform circle, 3 centered circle, 4 ring,
 * 5 rainbow ring, 6 color battle, 7 color wheel, 8 line, 9 spiral,
 * 10 rainbow spiral.  Every particle gets a position from its type, then
 * wrap or clamp (Physics.ensurePosition), and velocity 0 (Physics.setPosition,
 * Physics.java:297-303); types are kept.  The reference's setters draw from
 * unseeded java.util.Random; here particle k of the current order takes
 * output k of SplitMix64 streams of `seed` (uniform: the generator's 2^-24
 * grid, as plife_generate_particles; others: 53-bit uniforms and Box-Muller
 * Gaussians).  3D worlds take z uniform in [0, 1). */
int plife_set_positions(plife_physics *h, int32_t preset, uint64_t seed);

/* Physics.setTypes (Physics.java:509-511) with one of the app's type presets
 * (app/TypeSetterProvider.java:12-45), numbered in the app's order: 0 random,
 * 1 randomize 10%, 2 slices, 3 onion, 4 rotate, 5 flip, 6 more of first,
 * 7 kill still.  Draws as plife_set_positions. */
int plife_set_types(plife_physics *h, int32_t preset, uint64_t seed);

/* The app's matrix presets (app/MatrixGeneratorProvider.java:14-79), numbered
 * in its order: 0 random, 1 symmetry, 2 chains, 3 chains 2, 4 chains 3,
 * 5 snakes, 6 zero.  Random entries are 2u - 1 (DefaultMatrix.randomize,
 * DefaultMatrix.java:24-30) with u on the 2^-24 grid of SplitMix64 stream 5
 * of `seed`, in row order.  Host only (no handle, no device): writes
 * size * size doubles to out, row = acted-on type, size in [1, 256]. */
int plife_generate_matrix(int32_t preset, int32_t size, uint64_t seed, double *out);

/* Physics.update() (Physics.java:112-134) applied nsteps times.  Enqueues the
 * steps on the handle's stream; plife_synchronize or a download waits for
 * them.  To keep a stop prompt, it keeps only about 20 ms of device work
 * queued (at least 2 steps, at most 62), waiting for the oldest queued
 * launch before it queues more, and checks the stop flag before every launch
 * (one launch is 1 or 2 steps).  Returns PLIFE_STOPPED if
 * plife_force_update_stop() arrived: nothing more is queued, the steps
 * already queued complete, and the state is then at a step boundary;
 * plife_get_steps counts them. */
int plife_update(plife_physics *h, int32_t nsteps);

/* Physics.forceUpdateStop() (Physics.java:149-151).  The only call allowed
 * from another thread.  The reference stops before the next particle; here a
 * step is never cut short, and the stop takes effect before the next launch
 * (see plife_update).  The flag is cleared when plife_update starts, as
 * Physics.updateParticles sets its run flag (Physics.java:118). */
int plife_force_update_stop(plife_physics *h);

/* Steps taken by the handle's world since it was created: every step
 * plife_update has queued (a queued step always completes).  After
 * PLIFE_STOPPED and plife_synchronize this says where the world stands.
 * No counterpart in the reference, whose update() returns only when done. */
int plife_get_steps(plife_physics *h, int64_t *steps);

int plife_synchronize(plife_physics *h);

/* Thread-local message of the last failure on this thread ("" if none). */
const char *plife_last_error(void);

/* Parity exports of the last step (Physics.java:309-354): cell key of each
 * particle in the pre-sort order, the reference's containers array (end
 * offsets, ncells = nx^dim), and the permutation (sorted slot -> pre-sort
 * index).  Any pointer may be NULL.  ncells must equal the grid's cell count. */
int plife_get_cells(plife_physics *h, int32_t *keys, int32_t *cell_end, int32_t *perm,
                    int32_t ncells);

/* Cells per axis of the current settings: (int)floor(1/rmax). */
int plife_get_grid(plife_physics *h, int32_t *nx, int32_t *ncells);

/* Interacting ordered pairs (0 < d^2 <= rmax^2) counted in the last step. */
int plife_get_pair_count(plife_physics *h, int64_t *pairs);

/* Debug mode: when on, each step also records per particle (sorted slot) the
 * number of interacting neighbours and a 64-bit sum of a mix of their ids;
 * plife_get_neighbor_summary downloads them.  Off by default. */
int plife_set_debug(plife_physics *h, int32_t on);
int plife_get_neighbor_summary(plife_physics *h, int32_t n, int32_t *count, uint64_t *hash);

/* Debug mode: the acceleration each particle (sorted slot) received in the
 * last step -- the sum over its interacting neighbours of the accelerator's
 * output times rmax * force (Physics.java:437 without the dt), i.e.
 * (v_new - friction^(60 dt) v_old) / dt before the velocity is stored in
 * FP32 -- as FP32 (x, y, z) triples, acc[3n]. */
int plife_get_accelerations(plife_physics *h, int32_t n, float *acc);

/* Launch shape (no counterpart in the reference, which has no device).  Zero
 * fields mean automatic.  block_homes in {0, 64, 128, 256}: home particles
 * (threads) per force workgroup; list_cap in [0, 512]: per-thread LDS
 * candidate-list entries of the force kernel; stage_slack in [0, 16]: LDS
 * staging capacity as a multiple of the expected staged particles;
 * direct_launch 1: launch every kernel directly instead of replaying a HIP
 * graph of two steps; sparse_worlds: which worlds the sparse-world force
 * kernel (one thread per particle, stencil read from global memory) steps
 * instead of the staged one -- 0 those whose cells hold too few particles
 * for staging to pay, 1 none, 2 all; windowed_only 1: the staged force
 * kernel always screens homes in windows with per-thread candidate lists,
 * never on its balanced path (one window, homes dealt to lanes by their
 * kept-candidate counts); small_step: whether worlds that fit one
 * workgroup's LDS (up to about 10,000 particles in 2D, 7,000 in 3D) step in
 * one launch that bins, sorts and steps them in LDS -- 0 when every world of
 * the handle fits and no other launch-shape field is set, 1 never, 2
 * whenever every world fits (block_homes then sets the homes each of its
 * workgroups steps).  Results do not depend on the tuning.  Automatic
 * fields follow the world's clustering: every step measures the mean size of
 * the cell a particle sits in, and each plife_update plans its launches from
 * the latest measurement that has reached the host.  Two automatic choices
 * are measured instead, on the first update of at least 24 steps: the step
 * path (small_step 0, where every world fits) and, on the multi-launch path,
 * block_homes (256 or 128 homes per workgroup; kept per class of layout:
 * dimension, clustering, mean particles per cell, particle count).  Both
 * time the candidates on steps of that update, which the update counts. */
typedef struct {
    int32_t block_homes;
    int32_t list_cap;
    int32_t direct_launch;
    double stage_slack;
    int32_t sparse_worlds;
    int32_t windowed_only;
    int32_t small_step;
} plife_tuning;

int plife_set_tuning(plife_physics *h, const plife_tuning *t);
/* Current tuning with automatic fields resolved, and the LDS bytes per
 * workgroup of the last force launch. */
int plife_get_tuning(plife_physics *h, plife_tuning *t, int32_t *lds_bytes);

/* One Physics.update() (Physics.java:112-134) of one world whose state lives
 * in caller-owned device memory, enqueued on `stream` (a hipStream_t; NULL =
 * the default stream), that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It runs
 * plife_update's kernels, with every buffer they need beyond the state in a
 * static device array of the library (19.1 MB, held by each GPU the
 * library runs on until the process ends); it calls no hipMalloc, hipFree
 * or memset/memcpy, only kernel launches.
 *   - State in the device layout (plife_update's): positions as 32-bit
 *     fixed point (x = k * 2^-32 of the unit box), velocities FP32, types and
 *     ids int32, SoA arrays of n; pz, vz, opz, ovz only in 3D (else NULL).
 *     Inputs in any order; outputs in the new cell-sorted order (the
 *     reference's particle array after makeContainers' swap), ids carried.
 *     Inputs and outputs must not overlap.  Positions must already lie in
 *     the box (as plife_set_particles leaves them).
 *   - matrix: ntypes * ntypes FP64 in device memory, row = acted-on type.
 *   - pairs: optional device uint64 receiving the interacting pair count.
 *   - Limits: n <= 262144, nx^dim <= 2^20 cells (nx = floor(1/rmax));
 *     larger worlds return PLIFE_ERR_ARGUMENT.
 *   - Calls that overlap on one GPU share its array and race: serialise
 *     them (one stream, or events).  Each call clears what must start at zero.
 * Same results as plife_update stepping the same state, bit for bit. */
int plife_step_noalloc(const plife_settings *s, int32_t ntypes, const double *matrix, int32_t n,
                       const uint32_t *px, const uint32_t *py, const uint32_t *pz, const float *vx,
                       const float *vy, const float *vz, const int32_t *type, const int32_t *id,
                       uint32_t *opx, uint32_t *opy, uint32_t *opz, float *ovx, float *ovy,
                       float *ovz, int32_t *otype, int32_t *oid, uint64_t *pairs, void *stream);

/* Runs nsteps steps with HIP events around each phase and fills *out. */
int plife_profile(plife_physics *h, int32_t nsteps, plife_timing *out);

/* Runs nsteps steps and returns the device time between two events on the
 * handle's stream bracketing them (ms). */
int plife_time_steps(plife_physics *h, int32_t nsteps, double *ms);

/* ------------------------------------------------------------------------
 * Batches of independent worlds.  No counterpart in the reference, whose app
 * steps one world per Physics object (Physics.java:112-134): a batch steps
 * many worlds, each with its own settings, matrix and particle count, in the
 * same launches.  Each world's step is the single-world step -- its results
 * are bit-identical to stepping it alone in a plife_physics handle, whatever
 * else the batch holds.  All worlds of a batch share dim.  Worlds join and
 * leave between updates (continuous batching); ids are stable.  Limits:
 * 65535 worlds, 2^31 - 1 particles, 2^28 cells in total.
 * ------------------------------------------------------------------------ */
typedef struct plife_batch plife_batch;

int plife_batch_create(int32_t device, plife_batch **out);
int plife_batch_destroy(plife_batch *b);

/* Adds a world (settings, matrix and particles as for plife_set_settings and
 * plife_set_particles; copied in) and returns its id.  Ids of its particles
 * are 0..n-1 in the given order. */
int plife_batch_add_world(plife_batch *b, const plife_settings *s, int32_t ntypes,
                          const double *matrix, int32_t n, const double *pos, const double *vel,
                          const int32_t *types, int32_t *world_id);
/* plife_set_particle_slots / plife_set_positions / plife_set_types for one
 * world of the batch (slots of the world's own order, as plife_batch_get_world
 * returns it). */
int plife_batch_set_world_slots(plife_batch *b, int32_t world_id, int32_t k, const int32_t *slots,
                                const double *pos, const double *vel, const int32_t *types);
int plife_batch_set_world_positions(plife_batch *b, int32_t world_id, int32_t preset, uint64_t seed);
int plife_batch_set_world_types(plife_batch *b, int32_t world_id, int32_t preset, uint64_t seed);

/* Adds a world whose particles are generated on the device from `seed`, as
 * plife_generate_particles does. */
int plife_batch_add_generated_world(plife_batch *b, const plife_settings *s, int32_t ntypes,
                                    const double *matrix, int32_t n, uint64_t seed,
                                    int32_t *world_id);
/* Removes a world; its slots are reclaimed at the next update. */
int plife_batch_remove_world(plife_batch *b, int32_t world_id);
/* New settings / matrix for one world (dim fixed; the matrix may grow). */
int plife_batch_set_world_settings(plife_batch *b, int32_t world_id, const plife_settings *s,
                                   int32_t ntypes, const double *matrix);

/* nsteps steps of every world, all enqueued at once (unlike plife_update,
 * which bounds its queue: a batch's updates are short quanta, and batches on
 * their own streams overlap only if neither waits).  A stop request is
 * checked before every launch of this call. */
int plife_batch_update(plife_batch *b, int32_t nsteps);
int plife_batch_force_update_stop(plife_batch *b);
int plife_batch_synchronize(plife_batch *b);

/* One world's state in its current (cell-sorted) order; see plife_get_particles. */
int plife_batch_get_world(plife_batch *b, int32_t world_id, int32_t n, double *pos, double *vel,
                          int32_t *types, int32_t *ids);
/* Particle count, interacting pairs of the last step and steps taken. */
int plife_batch_world_info(plife_batch *b, int32_t world_id, int32_t *n, int64_t *pairs,
                           int64_t *steps);
int plife_batch_count(plife_batch *b, int32_t *nworlds, int64_t *nparticles);

int plife_batch_set_tuning(plife_batch *b, const plife_tuning *t);
int plife_batch_get_tuning(plife_batch *b, plife_tuning *t, int32_t *lds_bytes);
int plife_batch_time_steps(plife_batch *b, int32_t nsteps, double *ms);
int plife_batch_profile(plife_batch *b, int32_t nsteps, plife_timing *out);

#ifdef __cplusplus
}
#endif
#endif
