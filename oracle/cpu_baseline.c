/*
 * The CPU baseline (test and benchmark infrastructure only; nothing in the
 * product library links it): a multi-threaded restatement of the reference's
 * native matcher, structured as stereo_match's Matcher is
 * (native/stereo_match/src/lib.rs:107-381), timed beside the GPU path.
 *
 *  - rows in parallel: census, matching costs and both horizontal paths
 *    (lib.rs:147-168, 389-425; census.rs:18-84; paths.rs:85-126);
 *  - a down sweep over the vertical and two diagonal paths, column chunks in
 *    parallel per row (lib.rs:289-322; paths.rs:138-220);
 *  - an up sweep over the other three, with selection (lib.rs:323-349;
 *    select.rs:23-46);
 *  - rows in parallel: left-right check, then the 3x3 median
 *    (lib.rs:354-381; check.rs:7-85).
 *
 * Path states are bytes between two 255 sentinels, as paths.rs:1-8 lays
 * them out; the sentinels stand for libSGM's p2 - p1 padding, since both
 * only ever lose to the p2 jump. With 3 paths the down sweep follows the
 * vertical path alone and selects.
 *
 * Build: gcc -O3 -march=native -fopenmp -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define SENTINEL 255u

typedef struct {
    uint8_t *costs;      /* W*H*D */
    uint8_t *horizontal; /* W*H*D, sum of both horizontal paths */
    uint16_t *totals;    /* W*H*D */
    float *selected, *checked;
    uint32_t *keys;
    uint8_t *states;    /* 3 paths x 2 parities x (W+2) slots x (D+2) */
    uint8_t *minimums;  /* 3 x 2 x (W+2) */
    size_t volume_cap, pixel_cap, state_cap, minimum_cap;
} cpu_buffers;

static cpu_buffers g;

static int ensure(size_t volume, size_t pixels, size_t state_slots, int D) {
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
    size_t state_bytes = state_slots * (D + 2);
    if (state_bytes > g.state_cap) {
        free(g.states);
        g.states = (uint8_t *)malloc(state_bytes);
        if (!g.states) return -1;
        g.state_cap = state_bytes;
    }
    if (state_slots > g.minimum_cap) {
        free(g.minimums);
        g.minimums = (uint8_t *)malloc(state_slots);
        if (!g.minimums) return -1;
        g.minimum_cap = state_slots;
    }
    return 0;
}

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

static int cmp_i32(int32_t a, int32_t b) { return a < b; }

int cpu_sgm(const uint8_t *left, const uint8_t *right, int width, int height, int D, int paths,
            int p1i, int p2i, float uniqueness, float *out, int threads) {
    size_t pixels = (size_t)width * height, volume = pixels * D;
    size_t slots = (size_t)(width + 2);
    if (ensure(volume, pixels, 3 * 2 * slots, D)) return -1;
    uint8_t p1 = (uint8_t)p1i, p2 = (uint8_t)p2i;
    if (threads < 1) threads = omp_get_max_threads();
    const int n_sweep = paths == 8 ? 3 : 1;
    int offsets[3] = {0, 1, -1};
    size_t stride = (size_t)D + 2;

#pragma omp parallel num_threads(threads)
    {
        uint64_t *fl = (uint64_t *)malloc(8 * width), *fr = (uint64_t *)malloc(8 * width);
        uint8_t *backward = (uint8_t *)malloc((size_t)width * D);
        uint8_t states[4][258];

        /* rows: census, costs, horizontal paths */
#pragma omp for schedule(dynamic, 4)
        for (int y = 0; y < height; y++) {
            census_row(left, width, height, y, fl);
            census_row(right, width, height, y, fr);
            uint8_t *costs = g.costs + (size_t)y * width * D;
            for (int x = 0; x < width; x++)
                for (int d = 0; d < D; d++)
                    costs[(size_t)x * D + d] =
                        (uint8_t)__builtin_popcountll(fl[x] ^ (x >= d ? fr[x - d] : 0));
            uint8_t *sums = g.horizontal + (size_t)y * width * D;
            for (int s = 0; s < 4; s++) reset_state(states[s], D);
            uint8_t *fp = states[0], *fn = states[1], *bp = states[2], *bn = states[3];
            uint8_t fm = 0, bm = 0;
            for (int x = 0; x < width; x++) {
                int bx = width - 1 - x;
                fm = step(fp, fm, costs + (size_t)x * D, fn, D, p1, p2);
                bm = step(bp, bm, costs + (size_t)bx * D, bn, D, p1, p2);
                memcpy(sums + (size_t)x * D, fn + 1, D);
                memcpy(backward + (size_t)bx * D, bn + 1, D);
                uint8_t *t = fp;
                fp = fn;
                fn = t;
                t = bp;
                bp = bn;
                bn = t;
            }
            for (size_t i = 0; i < (size_t)width * D; i++) sums[i] = (uint8_t)(sums[i] + backward[i]);
        }

        /* two sweeps: down (totals), then up (select); 3 paths: down + select */
        for (int sweep = 0; sweep < (paths == 8 ? 2 : 1); sweep++) {
            int down = sweep == 0;
            int last = paths == 8 ? !down : 1;
#pragma omp for
            for (size_t s = 0; s < 3 * 2 * slots; s++) {
                reset_state(g.states + s * stride, D);
                g.minimums[s] = 0;
            }
            for (int row = 0; row < height; row++) {
                int y = down ? row : height - 1 - row;
                int cur = row % 2, prv = 1 - cur;
#pragma omp for schedule(static)
                for (int x = 0; x < width; x++) {
                    size_t at = ((size_t)y * width + x) * D;
                    const uint8_t *costs = g.costs + at;
                    uint16_t sum[256];
                    if (paths == 8 && !down)
                        for (int d = 0; d < D; d++) sum[d] = g.totals[at + d];
                    else
                        for (int d = 0; d < D; d++) sum[d] = g.horizontal[at + d];
                    for (int r = 0; r < n_sweep; r++) {
                        size_t source = ((size_t)r * 2 + prv) * slots + (size_t)(x + 1 - offsets[r]);
                        size_t target = ((size_t)r * 2 + cur) * slots + (size_t)(x + 1);
                        g.minimums[target] =
                            step(g.states + source * stride, g.minimums[source], costs,
                                 g.states + target * stride, D, p1, p2);
                        const uint8_t *v = g.states + target * stride + 1;
                        for (int d = 0; d < D; d++) sum[d] = (uint16_t)(sum[d] + v[d]);
                    }
                    if (last)
                        select_pixel(sum, D, x, uniqueness, g.selected + (size_t)y * width + x,
                                     g.keys + (size_t)y * width + x);
                    else
                        memcpy(g.totals + at, sum, 2 * (size_t)D);
                }
            }
        }

        /* rows: left-right check (check.rs:7-47) */
        uint32_t *rmap = (uint32_t *)malloc(4 * (size_t)width);
#pragma omp for schedule(dynamic, 4)
        for (int y = 0; y < height; y++) {
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
        free(rmap);

        /* rows: 3x3 median over i32 bit patterns (check.rs:49-85), then > 0 else -1 */
#pragma omp for schedule(dynamic, 4)
        for (int y = 0; y < height; y++) {
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
                /* Devillard's 19 compare-exchanges (filters.rs:10-30) */
                static const int net[19][2] = {{1, 2}, {4, 5}, {7, 8}, {0, 1}, {3, 4}, {6, 7}, {1, 2},
                                               {4, 5}, {7, 8}, {0, 3}, {5, 8}, {4, 7}, {3, 6}, {1, 4},
                                               {2, 5}, {4, 7}, {4, 2}, {6, 4}, {4, 2}};
                for (int e = 0; e < 19; e++) {
                    int32_t a = v[net[e][0]], b = v[net[e][1]];
                    v[net[e][0]] = cmp_i32(a, b) ? a : b;
                    v[net[e][1]] = cmp_i32(a, b) ? b : a;
                }
                float m;
                memcpy(&m, &v[4], 4);
                out[(size_t)y * width + x] = m > 0.0f ? m : -1.0f;
            }
        }
        free(fl);
        free(fr);
        free(backward);
    }
    return 0;
}

int cpu_max_threads(void) { return omp_get_max_threads(); }
