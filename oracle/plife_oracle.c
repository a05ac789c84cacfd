/*
 * plife_oracle.c -- FP64 CPU restatement of particle-life's physics step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (tests/,
 * __graft_entry__.smoke) and the CPU baseline leg of bench.py.  The product
 * path never calls it.
 *
 * Semantics follow the reference file by file; citations are relative to
 * src/main/java/com/particle_life/ in tom-mohr/particle-life-app:
 *
 *   step            backend/Physics.java:112-134   bin -> velocity phase -> position phase (Jacobi)
 *   binning         backend/Physics.java:309-354   stable counting sort, end offsets, swap
 *   cell key        backend/Physics.java:362-375   trunc(x/cs), clamp nx -> nx-1
 *   cell wrap       backend/Physics.java:377-395   one +/-nx step
 *   velocity        backend/Physics.java:397-441   friction, 9-cell stencil, predicate, accelerate
 *   position        backend/Physics.java:443-450   x = v*dt + x, then ensurePosition (499-505)
 *   min-image       backend/Range.java:68-81       [-0.5, 0.5)
 *   wrap / clamp    backend/Range.java:40-57, 83-96
 *   accelerator     app/Main.java:275-280          beta-zone force law
 *   thread split    backend/LoadDistributor.java:37-66  chunks of ceil(n/T)
 *
 * Vector arithmetic in the reference is JOML 1.10.1 (not vendored).  This file
 * evaluates it the way JOML does without -Djoml.useMathFma: lengthSquared as
 * x*x + (y*y + z*z), div(s) as a multiply by 1/s, mulAdd(a, b) as this*a + b.
 * Those three JOML details are recalled, not read from source, and can move
 * only the last FP64 ulp.  Build with -ffp-contract=off so no FMA is fused.
 *
 * The 3D mode (dim == 3) is an extension: the reference forces z = 0
 * (Range.java:43,71,86).  It uses a 27-cell stencil ordered z, y, x (outer to
 * inner), key = cx + cy*nx + cz*nx*nx, and wraps/clamps/min-images z like x, y.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c11 */

#include "plife_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    double p[3];
    double v[3];
    int32_t type;
    int32_t id;
} plo_particle;

uint64_t plo_id_hash(uint32_t id)
{
    uint64_t z = (uint64_t)id + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Physics.java:82-85 (called from makeContainers at 312-313 with cs = rmax). */
int plo_cells_per_axis(double rmax)
{
    return (int)floor(1.0 / rmax);
}

/* Physics.java:401: Math.pow(friction, 60 * dt). */
double plo_friction_factor(double friction, double dt)
{
    return pow(friction, 60 * dt);
}

/* Java's (int) cast of a double: truncation toward zero, NaN -> 0, saturation. */
static int java_d2i(double v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int)(-2147483647 - 1);
    return (int)v;
}

/* Physics.java:362-375: truncated FP64 quotient, the last index clamped. */
static int cell_key(const double *p, double cs, int nx, int dim)
{
    int cx = java_d2i(p[0] / cs);
    int cy = java_d2i(p[1] / cs);
    if (cx == nx) cx = nx - 1;
    if (cy == nx) cy = nx - 1;
    int key = cx + cy * nx;
    if (dim == 3) {
        int cz = java_d2i(p[2] / cs);
        if (cz == nx) cz = nx - 1;
        key += cz * nx * nx;
    }
    return key;
}

/* Physics.java:377-395: a single wrap step. */
static int wrap_cell(int c, int n)
{
    if (c < 0) return c + n;
    if (c >= n) return c - n;
    return c;
}

/* Range.java:74-81. */
static double wrap_connection(double v)
{
    if (v < -0.5) return v + 1;
    if (v >= 0.5) return v - 1;
    return v;
}

/* Range.java:46-57: repeated +1 / -1.  Non-finite values are returned as they
 * are (the reference would loop forever on +-inf); magnitudes beyond 2^52,
 * where +-1 no longer changes the value, are reduced with fmod instead. */
static double wrap_unit(double v)
{
    if (!isfinite(v)) return v;
    if (fabs(v) > 4503599627370496.0) {
        v = fmod(v, 1.0);
    }
    if (v < 0) {
        do {
            v += 1;
        } while (v < 0);
        return v;
    }
    while (v >= 1) v -= 1;
    return v;
}

/* Range.java:89-96. */
static double clamp_unit(double v)
{
    if (v < 0) return 0;
    if (v > 1) return 1;
    return v;
}

/* Physics.java:499-505 with Range.wrap / Range.clamp (z forced to 0 in 2D). */
static void ensure_position(double *p, int wrap, int dim)
{
    if (wrap) {
        p[0] = wrap_unit(p[0]);
        p[1] = wrap_unit(p[1]);
        p[2] = dim == 3 ? wrap_unit(p[2]) : 0;
    } else {
        p[0] = clamp_unit(p[0]);
        p[1] = clamp_unit(p[1]);
        p[2] = dim == 3 ? clamp_unit(p[2]) : 0;
    }
}

/* app/Main.java:275-280, operating in place like the lambda's pos.mul(...). */
void plo_accelerate(double a, double beta, double *rel)
{
    double dist = sqrt(rel[0] * rel[0] + (rel[1] * rel[1] + rel[2] * rel[2]));
    double f = dist < beta ? (dist / beta - 1)
                           : a * (1 - fabs(1 + beta - 2 * dist) / (1 - beta));
    double s = f / dist;
    rel[0] *= s;
    rel[1] *= s;
    rel[2] *= s;
}

typedef struct {
    const plo_settings *s;
    plo_particle *parts;    /* sorted array (Physics.particles after the swap) */
    const int32_t *cell_end;
    int n, nx, dim;
    double cs;
    int32_t *nbr_count;
    uint64_t *nbr_hash;
} step_ctx;

/* Physics.java:397-441 for the particle at sorted slot i. */
static void update_velocity(const step_ctx *c, int i)
{
    const plo_settings *s = c->s;
    plo_particle *p = &c->parts[i];
    const int nx = c->nx, dim = c->dim;

    double ff = pow(s->friction, 60 * s->dt);
    p->v[0] *= ff;
    p->v[1] *= ff;
    p->v[2] *= ff;

    /* stencil origin: floor, NOT clamped like the key (Physics.java:404-405) */
    int cx0 = java_d2i(floor(p->p[0] / c->cs));
    int cy0 = java_d2i(floor(p->p[1] / c->cs));
    int cz0 = dim == 3 ? java_d2i(floor(p->p[2] / c->cs)) : 0;
    const int zlo = dim == 3 ? -1 : 0, zhi = dim == 3 ? 1 : 0;

    const double r2 = s->rmax * s->rmax;
    const double inv_rmax = 1.0 / s->rmax;
    const double scale = s->rmax * s->force * s->dt;
    int32_t count = 0;
    uint64_t hash = 0;

    for (int oz = zlo; oz <= zhi; oz++) {
        for (int oy = -1; oy <= 1; oy++) {
            for (int ox = -1; ox <= 1; ox++) {
                int cx = wrap_cell(cx0 + ox, nx);
                int cy = wrap_cell(cy0 + oy, nx);
                int cz = dim == 3 ? wrap_cell(cz0 + oz, nx) : 0;
                if (s->wrap) {
                    cx = wrap_cell(cx, nx);
                    cy = wrap_cell(cy, nx);
                    if (dim == 3) cz = wrap_cell(cz, nx);
                } else if (cx < 0 || cx >= nx || cy < 0 || cy >= nx ||
                           (dim == 3 && (cz < 0 || cz >= nx))) {
                    continue;
                }
                int ci = cx + cy * nx + cz * nx * nx;
                int start = ci == 0 ? 0 : c->cell_end[ci - 1];
                int stop = c->cell_end[ci];
                for (int j = start; j < stop; j++) {
                    if (i == j) continue;
                    const plo_particle *q = &c->parts[j];
                    /* Physics.java:460-470 with Range.wrapConnection */
                    double d[3] = {q->p[0] - p->p[0], q->p[1] - p->p[1], q->p[2] - p->p[2]};
                    if (s->wrap) {
                        d[0] = wrap_connection(d[0]);
                        d[1] = wrap_connection(d[1]);
                        d[2] = dim == 3 ? wrap_connection(d[2]) : 0;
                    }
                    double d2 = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
                    if (d2 != 0 && d2 <= r2) {
                        d[0] *= inv_rmax;
                        d[1] *= inv_rmax;
                        d[2] *= inv_rmax;
                        double a = s->matrix[p->type * s->ntypes + q->type];
                        plo_accelerate(a, s->beta, d);
                        p->v[0] = p->v[0] + d[0] * scale;
                        p->v[1] = p->v[1] + d[1] * scale;
                        p->v[2] = p->v[2] + d[2] * scale;
                        count++;
                        if (c->nbr_hash) hash += plo_id_hash((uint32_t)q->id);   /* checker only */
                    }
                }
            }
        }
    }
    if (c->nbr_count) c->nbr_count[i] = count;
    if (c->nbr_hash) c->nbr_hash[i] = hash;
}

/* Physics.java:443-450. */
static void update_position(const step_ctx *c, int i)
{
    plo_particle *p = &c->parts[i];
    const double dt = c->s->dt;
    p->p[0] = p->v[0] * dt + p->p[0];
    p->p[1] = p->v[1] * dt + p->p[1];
    p->p[2] = p->v[2] * dt + p->p[2];
    ensure_position(p->p, c->s->wrap, c->dim);
}

typedef struct {
    const step_ctx *c;
    int start, stop, phase;
} chunk_job;

static void *run_chunk(void *arg)
{
    const chunk_job *j = (const chunk_job *)arg;
    for (int i = j->start; i < j->stop; i++) {
        if (j->phase == 0) update_velocity(j->c, i);
        else update_position(j->c, i);
    }
    return NULL;
}

/* LoadDistributor.distributeLoadEvenly (LoadDistributor.java:37-66):
 * chunks of ceil(n/T), joined before returning. */
static void distribute(const step_ctx *c, int n, int nthreads, int phase)
{
    if (n <= 0) return;
    if (nthreads <= 1) {
        chunk_job j = {c, 0, n, phase};
        run_chunk(&j);
        return;
    }
    int len = (int)ceil(n / (double)nthreads);
    int nchunks = (n + len - 1) / len;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nchunks);
    chunk_job *jobs = (chunk_job *)malloc(sizeof(chunk_job) * (size_t)nchunks);
    for (int k = 0; k < nchunks; k++) {
        jobs[k].c = c;
        jobs[k].start = k * len;
        jobs[k].stop = (k + 1) * len < n ? (k + 1) * len : n;
        jobs[k].phase = phase;
        pthread_create(&th[k], NULL, run_chunk, &jobs[k]);
    }
    for (int k = 0; k < nchunks; k++) pthread_join(th[k], NULL);
    free(th);
    free(jobs);
}

int plo_step(const plo_settings *s, int n, double *pos, double *vel,
             int32_t *type, int32_t *id, int nthreads, plo_debug *dbg)
{
    if (!s || n < 0 || (n > 0 && (!pos || !vel || !type || !id))) return -1;
    if (s->dim != 2 && s->dim != 3) return -2;
    if (!(s->rmax > 0) || s->rmax > 1 || s->ntypes <= 0 || !s->matrix) return -3;

    const int dim = s->dim;
    const double cs = s->rmax;                       /* Physics.java:312 */
    const int nx = plo_cells_per_axis(cs);           /* Physics.java:313 */
    const long long ncells_ll = dim == 3 ? (long long)nx * nx * nx : (long long)nx * nx;
    if (nx < 1 || ncells_ll > 0x7fffffffLL) return -4;
    const int ncells = (int)ncells_ll;

    for (int i = 0; i < n; i++) {
        if (type[i] < 0 || type[i] >= s->ntypes) return -5;
    }

    plo_particle *parts = (plo_particle *)malloc(sizeof(plo_particle) * (size_t)(n > 0 ? n : 1));
    plo_particle *sorted = (plo_particle *)malloc(sizeof(plo_particle) * (size_t)(n > 0 ? n : 1));
    int32_t *cont = (int32_t *)calloc((size_t)ncells, sizeof(int32_t));
    int32_t *keys = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!parts || !sorted || !cont || !keys || !perm) {
        free(parts); free(sorted); free(cont); free(keys); free(perm);
        return -6;
    }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            parts[i].p[k] = pos[3 * i + k];
            parts[i].v[k] = vel[3 * i + k];
        }
        parts[i].type = type[i];
        parts[i].id = id[i];
    }

    struct timespec t_begin, t_end;
    clock_gettime(CLOCK_MONOTONIC, &t_begin);

    /* makeContainers (Physics.java:309-354) */
    for (int i = 0; i < n; i++) {
        keys[i] = cell_key(parts[i].p, cs, nx, dim);
        cont[keys[i]]++;
    }
    int offset = 0;
    for (int c = 0; c < ncells; c++) {
        int cap = cont[c];
        cont[c] = offset;
        offset += cap;
    }
    for (int i = 0; i < n; i++) {
        int slot = cont[keys[i]]++;
        sorted[slot] = parts[i];
        perm[slot] = i;
    }
    /* cont[] now holds end offsets, the array is swapped to the sorted one */

    step_ctx c;
    c.s = s;
    c.parts = sorted;
    c.cell_end = cont;
    c.n = n;
    c.nx = nx;
    c.dim = dim;
    c.cs = cs;
    c.nbr_count = dbg ? dbg->nbr_count : NULL;
    c.nbr_hash = dbg ? dbg->nbr_hash : NULL;
    int32_t *tmp_count = NULL;
    if (dbg && !c.nbr_count) {
        tmp_count = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        c.nbr_count = tmp_count;
    }

    /* Physics.java:122-131: velocity phase for all, then position phase */
    distribute(&c, n, nthreads, 0);
    distribute(&c, n, nthreads, 1);
    clock_gettime(CLOCK_MONOTONIC, &t_end);

    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] = sorted[i].p[k];
            vel[3 * i + k] = sorted[i].v[k];
        }
        type[i] = sorted[i].type;
        id[i] = sorted[i].id;
    }
    if (dbg) {
        if (dbg->keys) memcpy(dbg->keys, keys, sizeof(int32_t) * (size_t)n);
        if (dbg->cell_end) memcpy(dbg->cell_end, cont, sizeof(int32_t) * (size_t)ncells);
        if (dbg->perm) memcpy(dbg->perm, perm, sizeof(int32_t) * (size_t)n);
        int64_t total = 0;
        for (int i = 0; i < n; i++) total += c.nbr_count[i];
        dbg->pair_count = total;
        dbg->seconds = (double)(t_end.tv_sec - t_begin.tv_sec) +
                       1e-9 * (double)(t_end.tv_nsec - t_begin.tv_nsec);
    }
    free(tmp_count);
    free(parts);
    free(sorted);
    free(cont);
    free(keys);
    free(perm);
    return 0;
}

int plo_bruteforce_neighbors(const plo_settings *s, int n, const double *pos,
                             const int32_t *id, int32_t *count, uint64_t *hash)
{
    if (!s || n < 0 || (n > 0 && (!pos || !id || !count || !hash))) return -1;
    const double r2 = s->rmax * s->rmax;
    for (int i = 0; i < n; i++) {
        int32_t cnt = 0;
        uint64_t h = 0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = pos[3 * j + k] - pos[3 * i + k];
            if (s->wrap) {
                d[0] = wrap_connection(d[0]);
                d[1] = wrap_connection(d[1]);
                d[2] = s->dim == 3 ? wrap_connection(d[2]) : 0;
            }
            double d2 = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
            if (d2 != 0 && d2 <= r2) {
                cnt++;
                h += plo_id_hash((uint32_t)id[j]);
            }
        }
        count[i] = cnt;
        hash[i] = h;
    }
    return 0;
}
