/*
 * The CPU baseline (test and benchmark infrastructure only; nothing in the
 * product library links it): a multi-threaded restatement of the reference's
 * native CPU path, stereo_match, with its work split and scheduled as
 * stereo_match does it, timed beside the GPU path.
 *
 *  - Matcher::disparity (native/stereo_match/src/lib.rs:107-144):
 *     8 paths: rows taken four at a time from a queue (parallel.rs:50-113):
 *       census, matching costs and both horizontal paths (lib.rs:147-168,
 *       389-425); then a down sweep over the vertical and both down
 *       diagonals, and an up sweep over the other three with selection
 *       (lib.rs:289-350), each over column chunks, four per thread, a chunk
 *       claimed by one thread at a time for up to eight rows and doing a row
 *       once the chunks beside it have done the row before (paths.rs:150-219;
 *       parallel.rs:159-250), waiting threads spinning, then yielding, then
 *       sleeping (parallel.rs:256-281);
 *     3 paths: one pass down, threads producing rows' costs and horizontal
 *       paths into a ring of rows of about 1 MiB and following the vertical
 *       path down the column chunks over the rows ready, with selection
 *       (lib.rs:170-286, 385);
 *     the left-right check of rows from a queue, then each median row once
 *       its three checked rows are done (lib.rs:352-381; check.rs:7-85);
 *  - Rectification: the sampling map of each camera built once and kept
 *    while the camera stays the same (segment.rs:335-352; rectify.rs:34-59),
 *    rows of both images from a queue (rectify.rs:132-147);
 *  - SpeckleFilter and Depth: rows from a queue (filters.rs:222-268; depth.rs:12-44).
 *    PointCloud and Quality run on one thread, as stereo_match's do (depth.rs:58-89;
 *    measures.rs:125-154): the oracle's.
 *
 * The work is run by OpenMP's thread pool, kept from one call to the next as
 * stereo_match's pool is (parallel.rs:283-409). Path states are bytes between
 * two 255 sentinels, as paths.rs:1-8 lays them out; the sentinels stand for
 * libSGM's p2 - p1 padding, since both only ever lose to the p2 jump.
 *
 * Build: gcc -O3 -fopenmp -ffp-contract=off (oracle/Makefile).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SENTINEL 255u
#define ROWS_PER_TAKE 4     /* parallel.rs:50 */
#define ROWS_PER_CLAIM 8    /* parallel.rs:155 */
#define CHUNKS_PER_THREAD 4 /* parallel.rs:166 */
#define MIN_CHUNK_COLUMNS 16
#define RING_BYTES (1 << 20) /* lib.rs:385 */

/* ---- scheduling (parallel.rs) ---- */

/* Rows handed out ROWS_PER_TAKE at a time. */
typedef struct {
    atomic_int next;
    int rows;
} row_queue;

static void queue_init(row_queue *q, int rows) {
    atomic_init(&q->next, 0);
    q->rows = rows;
}

static int queue_take(row_queue *q, int *first, int *end) {
    int start = atomic_fetch_add_explicit(&q->next, ROWS_PER_TAKE, memory_order_relaxed);
    if (start >= q->rows) return 0;
    *first = start;
    *end = start + ROWS_PER_TAKE < q->rows ? start + ROWS_PER_TAKE : q->rows;
    return 1;
}

/* Spin, then yield, then sleep 50 us (parallel.rs:256-281). */
typedef struct {
    unsigned rounds;
} backoff;

static void backoff_wait(backoff *b) {
    if (b->rounds < 6) {
        for (unsigned i = 0; i < 1u << b->rounds; i++) __builtin_ia32_pause();
    } else if (b->rounds < 10) {
        sched_yield();
    } else {
        struct timespec t = {0, 50000};
        nanosleep(&t, NULL);
    }
    if (b->rounds < 1000) b->rounds++;
}

/* Column chunks, each done a row at a time by the thread holding it. */
typedef struct {
    int count;
    int *first; /* count + 1 column bounds */
    atomic_int *progress;
    atomic_int *claimed;
} chunks_t;

static int chunks_init(chunks_t *c, int width, int threads) {
    int count = threads * CHUNKS_PER_THREAD;
    if (count > width / MIN_CHUNK_COLUMNS) count = width / MIN_CHUNK_COLUMNS;
    if (count < 1) count = 1;
    c->count = count;
    c->first = (int *)malloc(sizeof(int) * (count + 1));
    c->progress = (atomic_int *)malloc(sizeof(atomic_int) * count);
    c->claimed = (atomic_int *)malloc(sizeof(atomic_int) * count);
    if (!c->first || !c->progress || !c->claimed) return -1;
    for (int k = 0; k <= count; k++) c->first[k] = (int)((long)width * k / count);
    for (int k = 0; k < count; k++) {
        atomic_init(&c->progress[k], 0);
        atomic_init(&c->claimed[k], 0);
    }
    return 0;
}

static void chunks_free(chunks_t *c) {
    free(c->first);
    free(c->progress);
    free((void *)c->claimed);
}

static int progress(const chunks_t *c, int k) {
    return atomic_load_explicit(&c->progress[k], memory_order_acquire);
}

static int min_progress(const chunks_t *c) {
    int m = progress(c, 0);
    for (int k = 1; k < c->count; k++) {
        int p = progress(c, k);
        if (p < m) m = p;
    }
    return m;
}

static int try_claim(chunks_t *c, int k) {
    int expected = 0;
    return atomic_compare_exchange_strong_explicit(&c->claimed[k], &expected, 1,
                                                   memory_order_acquire, memory_order_relaxed);
}

static void release_chunk(chunks_t *c, int k) {
    atomic_store_explicit(&c->claimed[k], 0, memory_order_release);
}

/* ---- buffers kept from one call to the next ---- */

typedef struct {
    uint8_t *costs;      /* W*H*D (8 paths), or the ring's rows (3 paths) */
    uint8_t *horizontal; /* the same: the sum of both horizontal paths */
    uint16_t *totals;    /* W*H*D: the down sweep's sums */
    float *selected, *checked;
    uint32_t *keys;
    uint8_t *states;    /* 3 paths x 2 parities x (W+2) slots x (D+2) */
    uint8_t *minimums;  /* 3 x 2 x (W+2) */
    atomic_int *done;   /* per row: checked */
    atomic_int *produced; /* per ring slot: the row it holds + 1 */
    size_t volume_cap, pixel_cap, state_cap, minimum_cap, rows_cap;
} cpu_buffers;

static cpu_buffers g;

static int grow(void **p, size_t *cap, size_t bytes) {
    if (bytes <= *cap) return 0;
    free(*p);
    *p = malloc(bytes);
    *cap = *p ? bytes : 0;
    return *p ? 0 : -1;
}

static int ensure(size_t volume, size_t pixels, size_t state_slots, int D, int rows) {
    if (volume > g.volume_cap) {
        free(g.costs);
        free(g.horizontal);
        free(g.totals);
        g.costs = (uint8_t *)malloc(volume);
        g.horizontal = (uint8_t *)malloc(volume);
        g.totals = (uint16_t *)malloc(2 * volume);
        if (!g.costs || !g.horizontal || !g.totals) return -1;
        g.volume_cap = volume;
    }
    if (pixels > g.pixel_cap) {
        free(g.selected);
        free(g.checked);
        free(g.keys);
        g.selected = (float *)malloc(4 * pixels);
        g.checked = (float *)malloc(4 * pixels);
        g.keys = (uint32_t *)malloc(4 * pixels);
        if (!g.selected || !g.checked || !g.keys) return -1;
        g.pixel_cap = pixels;
    }
    if (grow((void **)&g.states, &g.state_cap, state_slots * (D + 2))) return -1;
    if (grow((void **)&g.minimums, &g.minimum_cap, state_slots)) return -1;
    if ((size_t)rows > g.rows_cap) {
        free((void *)g.done);
        free((void *)g.produced);
        g.done = (atomic_int *)malloc(sizeof(atomic_int) * rows);
        g.produced = (atomic_int *)malloc(sizeof(atomic_int) * rows);
        if (!g.done || !g.produced) return -1;
        g.rows_cap = rows;
    }
    return 0;
}

/* ---- the matcher's pieces ---- */

static inline uint8_t min8(uint8_t a, uint8_t b) { return a < b ? a : b; }

/* paths.rs:19-42: the next state from the previous (sentinel-padded) one. */
static inline uint8_t step(const uint8_t *previous, uint8_t previous_minimum,
                           const uint8_t *costs, uint8_t *next, int D, uint8_t p1,
                           uint8_t p2) {
    uint8_t jump = (uint8_t)(previous_minimum + p2);
    uint8_t minimum = 255;
    for (int d = 0; d < D; d++) {
        unsigned nb = (unsigned)min8(previous[d], previous[d + 2]) + p1;
        uint8_t neighbour = nb > 255 ? 255 : (uint8_t)nb;
        uint8_t value =
            (uint8_t)(min8(min8(previous[d + 1], jump), neighbour) - previous_minimum + costs[d]);
        next[d + 1] = value;
        minimum = min8(minimum, value);
    }
    return minimum;
}

static void reset_state(uint8_t *state, int D) {
    state[0] = SENTINEL;
    memset(state + 1, 0, D);
    state[D + 1] = SENTINEL;
}

static void census_row(const uint8_t *image, int width, int height, int y, uint64_t *out) {
    for (int x = 0; x < width; x++) {
        int centre = image[y * width + x];
        uint64_t f = 0;
        int bit = 0;
        for (int dy = -3; dy <= 3; dy++) {
            if (dy == 0) continue;
            int sy = y + dy < 0 ? 0 : (y + dy >= height ? height - 1 : y + dy);
            const uint8_t *row = image + (size_t)sy * width;
            for (int dx = -4; dx <= 4; dx++) {
                if (dx == 0) continue;
                int sx = x + dx < 0 ? 0 : (x + dx >= width ? width - 1 : x + dx);
                f |= (uint64_t)(row[sx] + 1 < centre) << bit;
                bit++;
            }
        }
        out[x] = f;
    }
}

/* What a thread computes a row's costs and horizontal paths with (lib.rs:389-425). */
typedef struct {
    uint64_t *fl, *fr;
    uint8_t *backward;
    uint8_t states[4][258];
} row_scratch;

static int scratch_init(row_scratch *s, int width, int D) {
    s->fl = (uint64_t *)malloc(8 * (size_t)width);
    s->fr = (uint64_t *)malloc(8 * (size_t)width);
    s->backward = (uint8_t *)malloc((size_t)width * D);
    return s->fl && s->fr && s->backward ? 0 : -1;
}

static void scratch_free(row_scratch *s) {
    free(s->fl);
    free(s->fr);
    free(s->backward);
}

/* Row y's costs and the byte sum of its two horizontal paths (census.rs:18-84;
 * paths.rs:85-126). */
static void costs_and_paths(row_scratch *s, const uint8_t *left, const uint8_t *right, int width,
                            int height, int y, int D, uint8_t p1, uint8_t p2, uint8_t *costs,
                            uint8_t *sums) {
    census_row(left, width, height, y, s->fl);
    census_row(right, width, height, y, s->fr);
    for (int x = 0; x < width; x++)
        for (int d = 0; d < D; d++)
            costs[(size_t)x * D + d] =
                (uint8_t)__builtin_popcountll(s->fl[x] ^ (x >= d ? s->fr[x - d] : 0));
    for (int k = 0; k < 4; k++) reset_state(s->states[k], D);
    uint8_t *fp = s->states[0], *fn = s->states[1], *bp = s->states[2], *bn = s->states[3];
    uint8_t fm = 0, bm = 0;
    for (int x = 0; x < width; x++) {
        int bx = width - 1 - x;
        fm = step(fp, fm, costs + (size_t)x * D, fn, D, p1, p2);
        bm = step(bp, bm, costs + (size_t)bx * D, bn, D, p1, p2);
        memcpy(sums + (size_t)x * D, fn + 1, D);
        memcpy(s->backward + (size_t)bx * D, bn + 1, D);
        uint8_t *t = fp;
        fp = fn;
        fn = t;
        t = bp;
        bp = bn;
        bn = t;
    }
    for (size_t i = 0; i < (size_t)width * D; i++) sums[i] = (uint8_t)(sums[i] + s->backward[i]);
}

/* select.rs:23-46 over u16 sums */
static void select_pixel(const uint16_t *sums, int D, int x, float uniqueness, float *disparity,
                         uint32_t *key) {
    uint16_t values[256];
    uint16_t minimum = 0xFFFF;
    for (int d = 0; d < D; d++) {
        values[d] = (uint16_t)(sums[d] + (d > x ? 0x8000 : 0));
        if (values[d] < minimum) minimum = values[d];
    }
    int winner = 0;
    while (values[winner] != minimum) winner++;
    int below = winner > 0 ? winner - 1 : 0, above = winner + 1 < D ? winner + 1 : D - 1;
    uint16_t second = 0xFFFF;
    for (int d = 0; d < D; d++)
        if ((d < below || d > above) && values[d] < second) second = values[d];
    float lower = values[below], upper = values[above], minimum_f32 = minimum;
    int accepted = (float)second > minimum_f32 * uniqueness;
    float denominator = lower + upper - 2.0f * minimum_f32;
    int limit = D - 1 < x ? D - 1 : x;
    int refine = denominator > 0.0f && winner > 0 && winner < limit;
    float correction = (lower - upper) / (2.0f * (denominator > 1.0f ? denominator : 1.0f));
    correction = correction < -0.5f ? -0.5f : (correction > 0.5f ? 0.5f : correction);
    if (accepted) {
        *disparity = (float)winner + (refine ? correction : 0.0f);
        *key = ((uint32_t)minimum << 16) | (uint32_t)winner;
    } else {
        *disparity = -1.0f;
        *key = ORC_INVALID_KEY;
    }
}

/* ---- the 8-path form ---- */

typedef struct {
    int width, height, D, down;
    size_t slots, stride;
    uint8_t p1, p2;
    float uniqueness;
} sweep_args;

static const int OFFSETS[3] = {0, 1, -1}; /* straight, from the left, from the right */

/* Row `row` of the sweep over one chunk's columns (paths.rs:182-201): the down sweep adds
 * its three paths to the horizontal sum into the totals; the up sweep adds its three to the
 * totals and selects. */
static void sweep_row(const sweep_args *a, const chunks_t *c, int chunk, int row) {
    const int width = a->width, D = a->D;
    const int y = a->down ? row : a->height - 1 - row;
    const int cur = row % 2, prv = 1 - cur;
    uint16_t sum[256];
    for (int x = c->first[chunk]; x < c->first[chunk + 1]; x++) {
        size_t at = ((size_t)y * width + x) * D;
        const uint8_t *costs = g.costs + at;
        if (a->down)
            for (int d = 0; d < D; d++) sum[d] = g.horizontal[at + d];
        else
            for (int d = 0; d < D; d++) sum[d] = g.totals[at + d];
        for (int r = 0; r < 3; r++) {
            size_t source = ((size_t)r * 2 + prv) * a->slots + (size_t)(x + 1 - OFFSETS[r]);
            size_t target = ((size_t)r * 2 + cur) * a->slots + (size_t)(x + 1);
            g.minimums[target] = step(g.states + source * a->stride, g.minimums[source], costs,
                                      g.states + target * a->stride, D, a->p1, a->p2);
            const uint8_t *v = g.states + target * a->stride + 1;
            for (int d = 0; d < D; d++) sum[d] = (uint16_t)(sum[d] + v[d]);
        }
        if (a->down)
            memcpy(g.totals + at, sum, 2 * (size_t)D);
        else
            select_pixel(sum, D, x, a->uniqueness, g.selected + (size_t)y * width + x,
                         g.keys + (size_t)y * width + x);
    }
}

/* Sweep::run's work loop (paths.rs:203-218), on the calling thread of `threads`: passes over
 * the chunks from its own share, each free one whose neighbours have done the row before
 * claimed for up to ROWS_PER_CLAIM rows (parallel.rs:199-233). */
static void sweep_work(const sweep_args *a, chunks_t *c, int thread, int threads) {
    backoff b = {0};
    const int count = c->count, height = a->height;
    for (;;) {
        int progressed = 0;
        for (int offset = 0; offset < count; offset++) {
            const int k = (int)(((long)thread * count / threads + offset) % count);
            int row = progress(c, k);
#define READY(k, row)                                                                 \
    ((row) < height && ((k) == 0 || progress(c, (k) - 1) >= (row)) &&                  \
     ((k) == count - 1 || progress(c, (k) + 1) >= (row)))
            if (!READY(k, row) || !try_claim(c, k)) continue;
            const int first = progress(c, k);
            row = first;
            while (READY(k, row) && row - first < ROWS_PER_CLAIM) {
                sweep_row(a, c, k, row);
                row++;
                atomic_store_explicit(&c->progress[k], row, memory_order_release);
            }
#undef READY
            release_chunk(c, k);
            progressed |= row > first;
        }
        if (progressed)
            b.rounds = 0;
        else if (min_progress(c) == height)
            break;
        else
            backoff_wait(&b);
    }
}

/* ---- the 3-path form (lib.rs:170-286) ---- */

typedef struct {
    const uint8_t *left, *right;
    int width, height, D, ring_rows;
    size_t row_length, stride;
    uint8_t p1, p2;
    float uniqueness;
    atomic_int next_row;
} ring_args;

static void vertical_row(const ring_args *a, const chunks_t *c, int chunk, int y) {
    const int width = a->width, D = a->D;
    const size_t ring_row = (size_t)(y % a->ring_rows) * a->row_length;
    const int cur = y % 2, prv = 1 - cur;
    uint16_t sums[256];
    for (int x = c->first[chunk]; x < c->first[chunk + 1]; x++) {
        const size_t at = ring_row + (size_t)x * D;
        size_t source = (size_t)prv * width + x, target = (size_t)cur * width + x;
        g.minimums[target] = step(g.states + source * a->stride, g.minimums[source], g.costs + at,
                                  g.states + target * a->stride, D, a->p1, a->p2);
        const uint8_t *v = g.states + target * a->stride + 1;
        for (int d = 0; d < D; d++) sums[d] = (uint16_t)(g.horizontal[at + d] + v[d]);
        select_pixel(sums, D, x, a->uniqueness, g.selected + (size_t)y * width + x,
                     g.keys + (size_t)y * width + x);
    }
}

static void ring_work(ring_args *a, chunks_t *c, row_scratch *s, int thread, int threads) {
    backoff b = {0};
    const int count = c->count, height = a->height;
    for (;;) {
        int progressed = 0;
        for (int offset = 0; offset < count; offset++) {
            const int k = (int)(((long)thread * count / threads + offset) % count);
#define READY(row)                                                                            \
    ((row) < height &&                                                                        \
     atomic_load_explicit(&g.produced[(row) % a->ring_rows], memory_order_acquire) == (row) + 1)
            int row = progress(c, k);
            if (!READY(row) || !try_claim(c, k)) continue;
            const int first = progress(c, k);
            row = first;
            while (READY(row) && row - first < ROWS_PER_CLAIM) {
                vertical_row(a, c, k, row);
                row++;
                atomic_store_explicit(&c->progress[k], row, memory_order_release);
            }
#undef READY
            release_chunk(c, k);
            progressed |= row > first;
        }
        /* A row may go into the ring once every chunk is done with the row it replaces. */
        int row = atomic_load_explicit(&a->next_row, memory_order_acquire);
        if (row < height && row < min_progress(c) + a->ring_rows &&
            atomic_compare_exchange_strong_explicit(&a->next_row, &row, row + 1,
                                                    memory_order_acq_rel, memory_order_relaxed)) {
            const size_t slot = (size_t)(row % a->ring_rows) * a->row_length;
            costs_and_paths(s, a->left, a->right, a->width, height, row, a->D, a->p1, a->p2,
                            g.costs + slot, g.horizontal + slot);
            atomic_store_explicit(&g.produced[row % a->ring_rows], row + 1, memory_order_release);
            progressed = 1;
        }
        if (progressed)
            b.rounds = 0;
        else if (min_progress(c) == height)
            break;
        else
            backoff_wait(&b);
    }
}

/* ---- left-right check and median (lib.rs:352-381) ---- */

static void check_row(int width, int y, uint32_t *rmap) {
    const uint32_t *keys = g.keys + (size_t)y * width;
    const float *sel = g.selected + (size_t)y * width;
    float *chk = g.checked + (size_t)y * width;
    for (int x = 0; x < width; x++) rmap[x] = ORC_INVALID_KEY;
    for (int x = 0; x < width; x++) {
        uint32_t key = keys[x];
        if (key != ORC_INVALID_KEY) {
            uint32_t w = key & 0xFFFF;
            uint32_t candidate = (key & 0xFFFF0000u) | (65535u - w);
            uint32_t *t = &rmap[x - (int)w];
            if (candidate < *t) *t = candidate;
        }
    }
    for (int x = 0; x < width; x++) {
        float value = sel[x];
        chk[x] = -1.0f;
        if (value >= 0.0f) {
            long low = (long)value;
            long high = low + ((float)low != value);
            int dis[2];
            for (int k = 0; k < 2; k++) {
                long d = k ? high : low, rx = x - d;
                dis[k] = rx >= 0 && rx < width && rmap[rx] != ORC_INVALID_KEY &&
                         labs(65535 - (long)(rmap[rx] & 0xFFFF) - d) > 1;
            }
            if (!(dis[0] && dis[1])) chk[x] = value;
        }
    }
}

/* 3x3 median over i32 bit patterns (check.rs:49-85), then > 0 else -1 */
static void median_row(int width, int height, int y, float *out) {
    /* Devillard's 19 compare-exchanges (filters.rs:10-30) */
    static const int net[19][2] = {{1, 2}, {4, 5}, {7, 8}, {0, 1}, {3, 4}, {6, 7}, {1, 2},
                                   {4, 5}, {7, 8}, {0, 3}, {5, 8}, {4, 7}, {3, 6}, {1, 4},
                                   {2, 5}, {4, 7}, {4, 2}, {6, 4}, {4, 2}};
    for (int x = 0; x < width; x++) {
        int32_t v[9];
        int n = 0;
        for (int dy = -1; dy <= 1; dy++) {
            int yy = y + dy < 0 ? 0 : (y + dy >= height ? height - 1 : y + dy);
            for (int dx = -1; dx <= 1; dx++) {
                int xx = x + dx < 0 ? 0 : (x + dx >= width ? width - 1 : x + dx);
                memcpy(&v[n++], &g.checked[(size_t)yy * width + xx], 4);
            }
        }
        for (int e = 0; e < 19; e++) {
            int32_t a = v[net[e][0]], b = v[net[e][1]];
            v[net[e][0]] = a < b ? a : b;
            v[net[e][1]] = a < b ? b : a;
        }
        float m;
        memcpy(&m, &v[4], 4);
        out[(size_t)y * width + x] = m > 0.0f ? m : -1.0f;
    }
}

int cpu_sgm(const uint8_t *left, const uint8_t *right, int width, int height, int D, int paths,
            int p1i, int p2i, float uniqueness, float *out, int threads) {
    const size_t pixels = (size_t)width * height;
    const size_t slots = (size_t)(width + 2), stride = (size_t)D + 2;
    if (threads < 1) threads = omp_get_max_threads();
    const uint8_t p1 = (uint8_t)p1i, p2 = (uint8_t)p2i;
    const size_t row_length = (size_t)width * D;
    int ring_rows = (int)(RING_BYTES / (2 * row_length));
    ring_rows = ring_rows < 2 ? 2 : (ring_rows > height ? height : ring_rows);
    const size_t volume = paths == 8 ? pixels * D : (size_t)ring_rows * row_length;
    if (ensure(volume, pixels, paths == 8 ? 3 * 2 * slots : 2 * (size_t)width, D, height)) return -1;
    chunks_t c;
    if (chunks_init(&c, width, threads)) return -1;
    int failed = 0;
    row_queue rows, check_rows, median_rows;
    queue_init(&rows, height);
    queue_init(&check_rows, height);
    queue_init(&median_rows, height);
    sweep_args sa = {width, height, D, 1, slots, stride, p1, p2, uniqueness};
    ring_args ra;
    ra.left = left;
    ra.right = right;
    ra.width = width;
    ra.height = height;
    ra.D = D;
    ra.ring_rows = ring_rows;
    ra.row_length = row_length;
    ra.stride = stride;
    ra.p1 = p1;
    ra.p2 = p2;
    ra.uniqueness = uniqueness;
    atomic_init(&ra.next_row, 0);
    for (int y = 0; y < height; y++) atomic_init(&g.done[y], 0);
    for (int r = 0; r < ring_rows; r++) atomic_init(&g.produced[r], 0);
    const int sweep_threads = threads < c.count ? threads : c.count;

#pragma omp parallel num_threads(threads)
    {
        const int thread = omp_get_thread_num();
        row_scratch s;
        uint32_t *rmap = (uint32_t *)malloc(4 * (size_t)width);
        if (scratch_init(&s, width, D) || !rmap) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp barrier
        if (!failed && paths == 8) {
            /* lib.rs:147-168: rows' costs and horizontal paths */
            int first, end;
            while (queue_take(&rows, &first, &end))
                for (int y = first; y < end; y++)
                    costs_and_paths(&s, left, right, width, height, y, D, p1, p2,
                                    g.costs + (size_t)y * row_length,
                                    g.horizontal + (size_t)y * row_length);
#pragma omp for
            for (size_t k = 0; k < 3 * 2 * slots; k++) {
                reset_state(g.states + k * stride, D);
                g.minimums[k] = 0;
            }
            /* lib.rs:289-350: down, then up with selection (threads beyond the chunks idle) */
            if (thread < sweep_threads) sweep_work(&sa, &c, thread, sweep_threads);
#pragma omp barrier
#pragma omp single
            {
                sa.down = 0;
                for (int k = 0; k < c.count; k++) atomic_store(&c.progress[k], 0);
            }
#pragma omp for
            for (size_t k = 0; k < 3 * 2 * slots; k++) {
                reset_state(g.states + k * stride, D);
                g.minimums[k] = 0;
            }
            if (thread < sweep_threads) sweep_work(&sa, &c, thread, sweep_threads);
#pragma omp barrier
        } else if (!failed) {
#pragma omp for
            for (size_t k = 0; k < 2 * (size_t)width; k++) {
                reset_state(g.states + k * stride, D);
                g.minimums[k] = 0;
            }
            ring_work(&ra, &c, &s, thread, threads);
#pragma omp barrier
        }
        if (!failed) {
            /* lib.rs:361-379: checked rows, then each median row once its three are checked */
            int first, end;
            while (queue_take(&check_rows, &first, &end))
                for (int y = first; y < end; y++) {
                    check_row(width, y, rmap);
                    atomic_store_explicit(&g.done[y], 1, memory_order_release);
                }
            while (queue_take(&median_rows, &first, &end))
                for (int y = first; y < end; y++) {
                    int up = y > 0 ? y - 1 : 0, down = y + 1 < height ? y + 1 : height - 1;
                    backoff b = {0};
                    while (!atomic_load_explicit(&g.done[up], memory_order_acquire) ||
                           !atomic_load_explicit(&g.done[y], memory_order_acquire) ||
                           !atomic_load_explicit(&g.done[down], memory_order_acquire))
                        backoff_wait(&b);
                    median_row(width, height, y, out);
                }
        }
        scratch_free(&s);
        free(rmap);
    }
    chunks_free(&c);
    return failed ? -1 : 0;
}

/* ---- Rectification (rectify.rs:34-147; segment.rs:335-352) ---- */

/* rectify.rs:107-128: the projection, operation for operation. */
static void project(const float *k, const float *dist, const float *ipr, float u, float v,
                    float *mx, float *my) {
    float x = ipr[0] * u + ipr[1] * v + ipr[2];
    float y = ipr[3] * u + ipr[4] * v + ipr[5];
    float w = ipr[6] * u + ipr[7] * v + ipr[8];
    x = x / w;
    y = y / w;
    float k1 = dist[0], k2 = dist[1], p1 = dist[2], p2 = dist[3], k3 = dist[4];
    float r2 = x * x + y * y;
    float radial = 1.0f + (k3 * r2 * r2 * r2 + (k2 * r2 * r2 + k1 * r2));
    float dx = x * radial + 2.0f * p1 * x * y + p2 * (r2 + 2.0f * x * x);
    float dy = y * radial + p1 * (r2 + 2.0f * y * y) + 2.0f * p2 * x * y;
    *mx = k[0] * dx + k[1] * dy + k[2];
    *my = k[4] * dy + k[5];
}

/* A camera's sampling map (rectify.rs:34-59): each output pixel's top-left tap and
 * weights, kept while the camera and sizes stay the same. */
typedef struct {
    float k[9], d[5], ipr[9];
    int width, height, fw, fh;
    int *columns, *rows;
    float *column_weights, *row_weights;
    size_t cap;
} rect_map;

static rect_map g_maps[2];

static int map_for(rect_map *m, const float *k, const float *d, const float *ipr, int width,
                   int height, int fw, int fh) {
    if (m->columns && m->width == width && m->height == height && m->fw == fw && m->fh == fh &&
        !memcmp(m->k, k, sizeof m->k) && !memcmp(m->d, d, sizeof m->d) &&
        !memcmp(m->ipr, ipr, sizeof m->ipr))
        return 0;
    size_t pixels = (size_t)width * height;
    if (pixels > m->cap) {
        free(m->columns);
        free(m->rows);
        free(m->column_weights);
        free(m->row_weights);
        m->columns = (int *)malloc(4 * pixels);
        m->rows = (int *)malloc(4 * pixels);
        m->column_weights = (float *)malloc(4 * pixels);
        m->row_weights = (float *)malloc(4 * pixels);
        if (!m->columns || !m->rows || !m->column_weights || !m->row_weights) {
            m->cap = 0;
            m->columns = NULL;
            return -1;
        }
        m->cap = pixels;
    }
    float sx = (float)((double)fw / (double)width);
    float sy = (float)((double)fh / (double)height);
#pragma omp parallel for
    for (int v = 0; v < height; v++)
        for (int u = 0; u < width; u++) {
            float mx, my;
            project(k, d, ipr, (float)u, (float)v, &mx, &my);
            float source_x = (mx + 0.5f) * sx - 0.5f;
            float source_y = (my + 0.5f) * sy - 0.5f;
            float fx0 = floorf(source_x), fy0 = floorf(source_y);
            size_t i = (size_t)v * width + u;
            m->columns[i] = (int)fx0;
            m->rows[i] = (int)fy0;
            m->column_weights[i] = source_x - fx0;
            m->row_weights[i] = source_y - fy0;
        }
    memcpy(m->k, k, sizeof m->k);
    memcpy(m->d, d, sizeof m->d);
    memcpy(m->ipr, ipr, sizeof m->ipr);
    m->width = width;
    m->height = height;
    m->fw = fw;
    m->fh = fh;
    return 0;
}

/* rectify.rs:62-104: row v of the image rectified from `frame` through map `m`. */
static void rectify_row(const rect_map *m, const uint8_t *frame, int v, uint8_t *out) {
    const int fw = m->fw, fh = m->fh;
    for (int u = 0; u < m->width; u++) {
        size_t i = (size_t)v * m->width + u;
        int x0 = m->columns[i], y0 = m->rows[i];
        float cw = m->column_weights[i], rw = m->row_weights[i];
        float s[4];
        if (x0 >= 0 && x0 + 1 < fw && y0 >= 0 && y0 + 1 < fh) {
            const uint8_t *p = frame + (size_t)y0 * fw + x0;
            s[0] = p[0];
            s[1] = p[1];
            s[2] = p[fw];
            s[3] = p[fw + 1];
        } else {
            for (int t = 0; t < 4; t++) {
                int xx = x0 + (t & 1), yy = y0 + (t >> 1);
                s[t] = (xx >= 0 && xx < fw && yy >= 0 && yy < fh) ? (float)frame[(size_t)yy * fw + xx]
                                                                   : 0.0f;
            }
        }
        float top = (1.0f - cw) * s[0] + cw * s[1];
        float bottom = (1.0f - cw) * s[2] + cw * s[3];
        float value = (1.0f - rw) * top + rw * bottom;
        float r = roundf(value);
        if (r < 0.0f) r = 0.0f;
        if (r > 255.0f) r = 255.0f;
        out[(size_t)v * m->width + u] = (uint8_t)r;
    }
}

/* Both frames rectified, rows of both from a queue (rectify.rs:132-147); each camera's map
 * built on its first use and kept while the camera is the same. */
int cpu_rectify_pair(const uint8_t *left_frame, const uint8_t *right_frame, int fw, int fh,
                     const float *left_k, const float *left_d, const float *left_ipr,
                     const float *right_k, const float *right_d, const float *right_ipr, int width,
                     int height, uint8_t *left_out, uint8_t *right_out, int threads) {
    if (threads < 1) threads = omp_get_max_threads();
    if (map_for(&g_maps[0], left_k, left_d, left_ipr, width, height, fw, fh) ||
        map_for(&g_maps[1], right_k, right_d, right_ipr, width, height, fw, fh))
        return -1;
    row_queue q;
    queue_init(&q, height);
#pragma omp parallel num_threads(threads)
    {
        int first, end;
        while (queue_take(&q, &first, &end))
            for (int v = first; v < end; v++) {
                rectify_row(&g_maps[0], left_frame, v, left_out);
                rectify_row(&g_maps[1], right_frame, v, right_out);
            }
    }
    return 0;
}

/* ---- SpeckleFilter (filters.rs:222-268): rows from a queue ---- */

int cpu_despeckle(const float *disparity, int width, int height, float min_support,
                  float max_difference, int radius, float *out, int threads) {
    if (threads < 1) threads = omp_get_max_threads();
    const int stride = width + 2 * radius;
    float *padded = (float *)malloc(sizeof(float) * (size_t)stride * (height + 2 * radius));
    if (!padded) return -1;
    for (size_t i = 0; i < (size_t)stride * (height + 2 * radius); i++) padded[i] = -1.0f;
    for (int y = 0; y < height; y++)
        memcpy(padded + (size_t)(y + radius) * stride + radius, disparity + (size_t)y * width,
               sizeof(float) * width);
    row_queue q;
    queue_init(&q, height);
    int failed = 0;
#pragma omp parallel num_threads(threads)
    {
        unsigned *support = (unsigned *)malloc(sizeof(unsigned) * width);
        if (!support) {
#pragma omp atomic write
            failed = 1;
        } else {
            int first, end;
            while (queue_take(&q, &first, &end))
                for (int y = first; y < end; y++) {
                    memset(support, 0, sizeof(unsigned) * width);
                    const float *centre = disparity + (size_t)y * width;
                    for (int dy = 0; dy <= 2 * radius; dy++)
                        for (int dx = 0; dx <= 2 * radius; dx++) {
                            if (dy == radius && dx == radius) continue;
                            const float *n = padded + (size_t)(y + dy) * stride + dx;
                            for (int x = 0; x < width; x++)
                                support[x] += (n[x] > 0.0f) & (fabsf(n[x] - centre[x]) <= max_difference);
                        }
                    for (int x = 0; x < width; x++)
                        out[(size_t)y * width + x] =
                            (centre[x] > 0.0f && (float)support[x] >= min_support) ? centre[x] : -1.0f;
                }
            free(support);
        }
    }
    free(padded);
    return failed ? -1 : 0;
}

/* ---- Depth (depth.rs:12-44): rows from a queue ---- */

int cpu_depth(const float *disparity, int width, int height, float f, float b, uint16_t *mm,
              float *m, int threads) {
    if (threads < 1) threads = omp_get_max_threads();
    const float fb = f * b;
    row_queue q;
    queue_init(&q, height);
#pragma omp parallel num_threads(threads)
    {
        int first, end;
        while (queue_take(&q, &first, &end))
            for (size_t i = (size_t)first * width; i < (size_t)end * width; i++) {
                float d = disparity[i];
                float metres = d > 0.0f ? fb / d : 0.0f;
                float rounded = roundf(metres * 1000.0f);
                m[i] = metres;
                mm[i] = rounded > 65535.0f ? 0 : (uint16_t)(rounded > 0.0f ? rounded : 0.0f);
            }
    }
    return 0;
}

int cpu_max_threads(void) { return omp_get_max_threads(); }
