/*
 * sanitize_main.c -- exercises the FP64 oracle under AddressSanitizer +
 * UndefinedBehaviorSanitizer and under ThreadSanitizer (test infrastructure
 * only; built by `make -C oracle sanitize`, run by tests/test_oracle_sanitize.py).
 *
 * It steps small 2D and 3D worlds (periodic and clamped, coarse grids with
 * repeated stencil cells, an empty and a one-particle world) with 1 and 4
 * threads -- the threaded phases restate LoadDistributor.distributeLoadEvenly
 * (backend/LoadDistributor.java:37-66) -- and checks that every thread count
 * gives the same bits, and that the cell-list neighbour counts equal the
 * brute-force ones.  Exit status 0 means every check passed; the sanitizers
 * abort with a report on the first error they find.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "plife_oracle.h"

static uint64_t rng_state;

static double uniform(void)
{
    /* SplitMix64, top 53 bits */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53;
}

typedef struct {
    int n, dim, wrap, ntypes, steps;
    double rmax;
} world_case;

static int run_case(const world_case *wc)
{
    const int n = wc->n;
    const size_t n3 = (size_t)3 * (size_t)(n > 0 ? n : 1);
    double *matrix = malloc(sizeof(double) * (size_t)wc->ntypes * (size_t)wc->ntypes);
    for (int i = 0; i < wc->ntypes * wc->ntypes; i++) matrix[i] = 2.0 * uniform() - 1.0;
    plo_settings s = {wc->dim, wc->wrap, wc->rmax, 0.75, 1.0, 0.02, 0.3, wc->ntypes, matrix};

    double *pos[2], *vel[2];
    int32_t *type[2], *id[2];
    for (int k = 0; k < 2; k++) {
        pos[k] = calloc(n3, sizeof(double));
        vel[k] = calloc(n3, sizeof(double));
        type[k] = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
        id[k] = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    }
    for (int i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) pos[0][3 * i + d] = d < wc->dim ? uniform() : 0.0;
        type[0][i] = (int32_t)(uniform() * wc->ntypes);
        id[0][i] = i;
    }
    memcpy(pos[1], pos[0], n3 * sizeof(double));
    memcpy(type[1], type[0], sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    memcpy(id[1], id[0], sizeof(int32_t) * (size_t)(n > 0 ? n : 1));

    const int nx = plo_cells_per_axis(wc->rmax);
    const int ncells = wc->dim == 3 ? nx * nx * nx : nx * nx;
    int32_t *keys = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    int32_t *cell_end = calloc((size_t)ncells, sizeof(int32_t));
    int32_t *perm = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    int32_t *cnt = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    uint64_t *hash = calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));
    int32_t *bf_cnt = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
    uint64_t *bf_hash = calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));

    int bad = 0;
    for (int step = 0; step < wc->steps && !bad; step++) {
        /* neighbour sets of the state about to be stepped, by brute force */
        if (plo_bruteforce_neighbors(&s, n, pos[0], id[0], bf_cnt, bf_hash)) bad = 1;
        for (int k = 0; k < 2 && !bad; k++) {
            plo_debug dbg = {keys, cell_end, perm, cnt, hash, 0, 0.0};
            if (plo_step(&s, n, pos[k], vel[k], type[k], id[k], k == 0 ? 1 : 4, &dbg)) bad = 2;
            if (k == 0 && nx >= 3) {
                /* cell-list sets (sorted slots) equal the brute-force ones (pre-sort order) */
                for (int slot = 0; slot < n && !bad; slot++)
                    if (cnt[slot] != bf_cnt[perm[slot]] || hash[slot] != bf_hash[perm[slot]]) bad = 3;
            }
        }
        if (!bad && (memcmp(pos[0], pos[1], n3 * sizeof(double)) ||
                     memcmp(vel[0], vel[1], n3 * sizeof(double)) ||
                     memcmp(id[0], id[1], sizeof(int32_t) * (size_t)(n > 0 ? n : 1))))
            bad = 4;   /* thread count changed the result */
    }
    printf("case n=%d dim=%d wrap=%d rmax=%g: %s (%d)\n", wc->n, wc->dim, wc->wrap, wc->rmax,
           bad ? "FAILED" : "ok", bad);
    for (int k = 0; k < 2; k++) {
        free(pos[k]);
        free(vel[k]);
        free(type[k]);
        free(id[k]);
    }
    free(matrix);
    free(keys);
    free(cell_end);
    free(perm);
    free(cnt);
    free(hash);
    free(bf_cnt);
    free(bf_hash);
    return bad;
}

int main(void)
{
    const world_case cases[] = {
        {1500, 2, 1, 6, 3, 0.05},
        {1200, 2, 0, 5, 3, 0.06},
        {800, 3, 1, 4, 2, 0.12},
        {700, 3, 0, 4, 2, 0.15},
        {300, 2, 1, 3, 2, 0.4},    /* nx = 2: repeated stencil cells */
        {60, 3, 1, 3, 2, 0.9},     /* nx = 1 */
        {1, 2, 1, 2, 2, 0.1},
        {0, 2, 1, 2, 2, 0.1},
    };
    rng_state = 0x5EED5A7ull;
    int bad = 0;
    for (size_t c = 0; c < sizeof cases / sizeof cases[0]; c++) bad |= run_case(&cases[c]);
    printf(bad ? "FAILED\n" : "all ok\n");
    return bad ? 1 : 0;
}
