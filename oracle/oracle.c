/*
 * The oracle (test infrastructure only; see oracle.h). A scalar restatement
 * of NxStereo's semi-global matching and neighbour stages, written for
 * clarity over speed: full i32 path volumes, one direction at a time, as the
 * reference's own step-by-step transcription does
 * (native/stereo_match/src/reference.rs:1-2).
 *
 * f32 arithmetic is kept to one rounding per operation: build with
 * -ffp-contract=off and without -ffast-math.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* reference.rs:8-18: dy in -3..=3 (outer), dx in -4..=4 (inner), both non-zero. */
static void neighbours(int *dys, int *dxs) {
    int n = 0;
    for (int dy = -3; dy <= 3; dy++)
        for (int dx = -4; dx <= 4; dx++)
            if (dy != 0 && dx != 0) {
                dys[n] = dy;
                dxs[n] = dx;
                n++;
            }
}

/* reference.rs:20-38; nx_stereo.ex:491-509 */
void orc_census(const uint8_t *image, int width, int height, uint64_t *features) {
    int dys[48], dxs[48];
    neighbours(dys, dxs);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            int centre = image[y * width + x];
            uint64_t f = 0;
            for (int bit = 0; bit < 48; bit++) {
                int sy = clampi(y + dys[bit], 0, height - 1);
                int sx = clampi(x + dxs[bit], 0, width - 1);
                if ((int)image[sy * width + sx] + 1 < centre) f |= (uint64_t)1 << bit;
            }
            features[(size_t)y * width + x] = f;
        }
}

/* reference.rs:41-63; nx_stereo.ex:528-542 */
void orc_costs(const uint64_t *left, const uint64_t *right, int width, int height,
               int disparities, uint8_t *costs) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++)
            for (int d = 0; d < disparities; d++) {
                uint64_t r = x >= d ? right[(size_t)y * width + x - d] : 0;
                costs[((size_t)y * width + x) * disparities + d] =
                    (uint8_t)__builtin_popcountll(left[(size_t)y * width + x] ^ r);
            }
}

/* reference.rs:65-121; nx_stereo.ex:575-606. Each pixel continues the path
 * from (y - dy, x - dx), from zero where that lies outside the image. */
void orc_path(const uint8_t *costs, int width, int height, int D, int dy, int dx, int p1, int p2,
              int32_t *out) {
    int32_t *normalized = (int32_t *)malloc(sizeof(int32_t) * D);
    for (int yi = 0; yi < height; yi++) {
        int y = dy < 0 ? height - 1 - yi : yi;
        for (int xi = 0; xi < width; xi++) {
            int x = dx < 0 ? width - 1 - xi : xi;
            int py = y - dy, px = x - dx;
            int inside = py >= 0 && py < height && px >= 0 && px < width;
            const int32_t *previous = inside ? out + ((size_t)py * width + px) * D : NULL;
            int32_t minimum = 0;
            if (previous) {
                minimum = previous[0];
                for (int d = 1; d < D; d++)
                    if (previous[d] < minimum) minimum = previous[d];
            }
            for (int d = 0; d < D; d++) normalized[d] = previous ? previous[d] - minimum : 0;
            size_t at = ((size_t)y * width + x) * D;
            for (int d = 0; d < D; d++) {
                int32_t lower = d == 0 ? p2 - p1 : normalized[d - 1];
                int32_t upper = d == D - 1 ? p2 - p1 : normalized[d + 1];
                int32_t v = normalized[d];
                if (p2 < v) v = p2;
                int32_t n = (lower < upper ? lower : upper) + p1;
                if (n < v) v = n;
                out[at + d] = v + costs[at + d];
            }
        }
    }
    free(normalized);
}

/* reference.rs:123-146 */
void orc_aggregate(const uint8_t *costs, int width, int height, int D, int paths, int p1, int p2,
                   uint16_t *sums) {
    static const int directions[8][2] = {{0, 1}, {0, -1}, {1, 0},   {-1, 0},
                                         {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    size_t volume = (size_t)width * height * D;
    int32_t *path = (int32_t *)malloc(sizeof(int32_t) * volume);
    memset(sums, 0, sizeof(uint16_t) * volume);
    int count = paths == 8 ? 8 : 3;
    for (int r = 0; r < count; r++) {
        orc_path(costs, width, height, D, directions[r][0], directions[r][1], p1, p2, path);
        for (size_t i = 0; i < volume; i++) sums[i] = (uint16_t)(sums[i] + path[i]);
    }
    free(path);
}

/* reference.rs:148-186; nx_stereo.ex:608-641 */
void orc_select(const uint16_t *sums, int width, int height, int D, float uniqueness,
                float *disparity, uint32_t *keys) {
    int32_t *values = (int32_t *)malloc(sizeof(int32_t) * D);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            size_t at = ((size_t)y * width + x) * D;
            for (int d = 0; d < D; d++) values[d] = sums[at + d] + (d > x ? 0x8000 : 0);
            int32_t minimum = values[0];
            int winner = 0;
            for (int d = 1; d < D; d++)
                if (values[d] < minimum) {
                    minimum = values[d];
                    winner = d;
                }
            int32_t second = INT32_MAX;
            for (int d = 0; d < D; d++)
                if (abs(d - winner) > 1 && values[d] < second) second = values[d];
            float lower = (float)values[winner - 1 < 0 ? 0 : winner - 1];
            float upper = (float)values[winner + 1 > D - 1 ? D - 1 : winner + 1];
            float minimum_f32 = (float)minimum;
            int accepted = (float)second > minimum_f32 * uniqueness;
            float denominator = lower + upper - 2.0f * minimum_f32;
            float den1 = denominator > 1.0f ? denominator : 1.0f;
            float correction = (lower - upper) / (2.0f * den1);
            if (correction < -0.5f) correction = -0.5f;
            if (correction > 0.5f) correction = 0.5f;
            int limit = D - 1 < x ? D - 1 : x;
            int refine = denominator > 0.0f && winner > 0 && winner < limit;
            float value = (float)winner + (refine ? correction : 0.0f);
            uint32_t key = ((uint32_t)minimum << 16) | (uint32_t)winner;
            disparity[(size_t)y * width + x] = accepted ? value : -1.0f;
            keys[(size_t)y * width + x] = accepted ? key : ORC_INVALID_KEY;
        }
    free(values);
}

/* reference.rs:188-235; nx_stereo.ex:643-683. The right map is the stacked
 * minimum over disparities, as the Nx code and the transcription build it. */
void orc_consistency(const float *disparity, const uint32_t *keys, int width, int height, int D,
                     float *out) {
    uint32_t *right = (uint32_t *)malloc(sizeof(uint32_t) * width);
    for (int y = 0; y < height; y++) {
        const uint32_t *row = keys + (size_t)y * width;
        for (int rx = 0; rx < width; rx++) {
            uint32_t best = ORC_INVALID_KEY;
            for (int d = 0; d < D; d++) {
                uint32_t key = rx + d < width ? row[rx + d] : ORC_INVALID_KEY;
                uint32_t candidate = ORC_INVALID_KEY;
                if (key != ORC_INVALID_KEY && (int)(key & 0xFFFF) == d)
                    candidate = (key & 0xFFFF0000u) | (uint32_t)(65535 - d);
                if (candidate < best) best = candidate;
            }
            right[rx] = best;
        }
        for (int x = 0; x < width; x++) {
            float value = disparity[(size_t)y * width + x];
            int lo = (int)floorf(value), hi = (int)ceilf(value);
            int rejected = 1;
            for (int k = 0; k < 2; k++) {
                int d = k == 0 ? lo : hi;
                int rx = x - d;
                int inside = rx >= 0 && rx < width;
                uint32_t found = right[clampi(rx, 0, width - 1)];
                int right_d = 65535 - (int)(found & 0xFFFF);
                int disagrees = inside && found != ORC_INVALID_KEY && abs(right_d - d) > 1;
                rejected = rejected && disagrees;
            }
            out[(size_t)y * width + x] = (value >= 0.0f && !rejected) ? value : -1.0f;
        }
    }
    free(right);
}

/* f32 total order (f32::total_cmp); the maps hold -1.0 or non-negative values. */
static int total_cmp(const void *a, const void *b) {
    int32_t ia, ib;
    memcpy(&ia, a, 4);
    memcpy(&ib, b, 4);
    ia ^= (int32_t)(((uint32_t)(ia >> 31)) >> 1);
    ib ^= (int32_t)(((uint32_t)(ib >> 31)) >> 1);
    return (ia > ib) - (ia < ib);
}

/* reference.rs:237-254; image.ex:483-505 */
void orc_median3(const float *image, int width, int height, float *out) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            float values[9];
            int n = 0;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++)
                    values[n++] = image[(size_t)clampi(y + dy, 0, height - 1) * width +
                                        clampi(x + dx, 0, width - 1)];
            qsort(values, 9, sizeof(float), total_cmp);
            out[(size_t)y * width + x] = values[4];
        }
}

/* reference.rs:256-277; nx_stereo.ex:475-480: sgm/3's output, the median of the
 * checked map (a match at 0 stays 0). */
int orc_sgm_raw(const uint8_t *left, const uint8_t *right, int width, int height, int D,
                int paths, int p1, int p2, float uniqueness, float *disparity) {
    size_t pixels = (size_t)width * height, volume = pixels * D;
    uint64_t *fl = (uint64_t *)malloc(8 * pixels), *fr = (uint64_t *)malloc(8 * pixels);
    uint8_t *costs = (uint8_t *)malloc(volume);
    uint16_t *sums = (uint16_t *)malloc(2 * volume);
    float *selected = (float *)malloc(4 * pixels), *checked = (float *)malloc(4 * pixels);
    uint32_t *keys = (uint32_t *)malloc(4 * pixels);
    int ok = fl && fr && costs && sums && selected && checked && keys;
    if (ok) {
        orc_census(left, width, height, fl);
        orc_census(right, width, height, fr);
        orc_costs(fl, fr, width, height, D, costs);
        orc_aggregate(costs, width, height, D, paths, p1, p2, sums);
        orc_select(sums, width, height, D, uniqueness, selected, keys);
        orc_consistency(selected, keys, width, height, D, checked);
        orc_median3(checked, width, height, disparity);
    }
    free(fl);
    free(fr);
    free(costs);
    free(sums);
    free(selected);
    free(checked);
    free(keys);
    return ok ? 0 : -1;
}

/* stage/disparity.ex:162; segment.rs:407-418: a match at 0 is unmatched. */
static void positive_only(float *disparity, size_t pixels) {
    for (size_t i = 0; i < pixels; i++)
        if (!(disparity[i] > 0.0f)) disparity[i] = -1.0f;
}

int orc_sgm(const uint8_t *left, const uint8_t *right, int width, int height, int D, int paths,
            int p1, int p2, float uniqueness, float *disparity) {
    int rc = orc_sgm_raw(left, right, width, height, D, paths, p1, p2, uniqueness, disparity);
    if (rc == 0) positive_only(disparity, (size_t)width * height);
    return rc;
}

/* image.ex:458-476 (area_weights): each output's taps, input index and f32 weight share /
 * scale of the input span [o scale, (o + 1) scale), computed in f64 and rounded to f32
 * (parity unpinned for the sub-1e-6 contributions of those taps, which round away).
 * Returns the taps per output (padding with zero weights is left out: adding 0 changes no
 * sum). index/weight hold size * (ceil(scale) + 1) entries. */
static int area_taps(int size, int in_size, int *index, float *weight, int *count) {
    double scale = (double)in_size / (double)size;
    int most = (int)ceil(scale) + 1;
    for (int o = 0; o < size; o++) {
        double first = o * scale, last = (o + 1) * scale;
        int n = 0;
        for (int input = (int)floor(first); input <= (int)ceil(last) - 1; input++) {
            double hi = input + 1 < last ? input + 1 : last;
            double lo = input > first ? input : first;
            double share = hi - lo;
            if (share > 0) {
                /* (o + 1) scale can round past in_size in f64, giving a last tap one past
                 * the input with a weight near 1e-15; it changes no rounded output, so it
                 * reads the last input instead of past it. */
                index[o * most + n] = input < in_size ? input : in_size - 1;
                weight[o * most + n] = (float)(share / scale);
                n++;
            }
        }
        count[o] = n;
    }
    return most;
}

/* image.ex:402-456 (resize_area): the f32 area resize of a u8 frame to {height, width}:
 * rows first (each output row the weighted sum of its input rows, taps in order), then
 * columns the same way on that f32 image, then rounded half away from zero, clipped to
 * 0..255. The frame's own size copies it. */
void orc_resize_area(const uint8_t *frame, int fw, int fh, int width, int height,
                     uint8_t *out) {
    if (fw == width && fh == height) {
        memcpy(out, frame, (size_t)fw * fh);
        return;
    }
    int ry = (int)ceil((double)fh / height) + 1, rx = (int)ceil((double)fw / width) + 1;
    int *yi = (int *)malloc(sizeof(int) * height * ry), *yc = (int *)malloc(sizeof(int) * height);
    float *yw = (float *)malloc(sizeof(float) * height * ry);
    int *xi = (int *)malloc(sizeof(int) * width * rx), *xc = (int *)malloc(sizeof(int) * width);
    float *xw = (float *)malloc(sizeof(float) * width * rx);
    float *rows = (float *)malloc(sizeof(float) * (size_t)height * fw);
    area_taps(height, fh, yi, yw, yc);
    area_taps(width, fw, xi, xw, xc);
    for (int y = 0; y < height; y++)
        for (int c = 0; c < fw; c++) {
            float t = 0.0f;
            for (int k = 0; k < yc[y]; k++) {
                float product = yw[y * ry + k] * (float)frame[(size_t)yi[y * ry + k] * fw + c];
                t = k == 0 ? product : t + product;
            }
            rows[(size_t)y * fw + c] = t;
        }
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            float t = 0.0f;
            for (int k = 0; k < xc[x]; k++) {
                float product = xw[x * rx + k] * rows[(size_t)y * fw + xi[x * rx + k]];
                t = k == 0 ? product : t + product;
            }
            float r = roundf(t);
            out[(size_t)y * width + x] = (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
        }
    free(yi);
    free(yc);
    free(yw);
    free(xi);
    free(xc);
    free(xw);
    free(rows);
}

/* ---- Band search and coarse-to-fine (nx_stereo.ex:164-471) ---- */

/* nx_stereo.ex:236-247: the rounded 2x2 average, (sum + 2) / 4. */
void orc_shrink(const uint8_t *image, int width, int height, uint8_t *out) {
    int hw = width / 2, hh = height / 2;
    for (int y = 0; y < hh; y++)
        for (int x = 0; x < hw; x++) {
            const uint8_t *a = image + (size_t)(2 * y) * width + 2 * x;
            int sum = a[0] + a[1] + a[width] + a[width + 1];
            out[(size_t)y * hw + x] = (uint8_t)((sum + 2) / 4);
        }
}

/* nx_stereo.ex:490-525: bit k, row-major over the window's neighbours, is set when the
 * clamped neighbour + 1 < the anchor. sparse_9x7 (window 0): dy in -3..3, dx in -4..4,
 * both non-zero, anchor the centre (orc_census). dense_9x5 (1): dy in -2..2, dx in
 * -4..4 but the centre. dense_9x5_blend (2): dense_9x5 with the anchor
 * (sum of w(dy)w(dx) * neighbour over the 3x3, w = 1,2,1, + 16 * centre + 24) / 32. */
void orc_census_window(const uint8_t *image, int width, int height, int window,
                       uint64_t *features) {
    if (window == 0) {
        orc_census(image, width, height, features);
        return;
    }
    static const int weight[3] = {1, 2, 1};
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            int anchor = image[(size_t)y * width + x];
            if (window == 2) {
                int sum = 0;
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++)
                        sum += weight[dy + 1] * weight[dx + 1] *
                               image[(size_t)clampi(y + dy, 0, height - 1) * width +
                                     clampi(x + dx, 0, width - 1)];
                anchor = (sum + 16 * anchor + 24) / 32;
            }
            uint64_t f = 0;
            int bit = 0;
            for (int dy = -2; dy <= 2; dy++)
                for (int dx = -4; dx <= 4; dx++) {
                    if (dy == 0 && dx == 0) continue;
                    int v = image[(size_t)clampi(y + dy, 0, height - 1) * width +
                                  clampi(x + dx, 0, width - 1)];
                    if (v + 1 < anchor) f |= (uint64_t)1 << bit;
                    bit++;
                }
            features[(size_t)y * width + x] = f;
        }
}

/* nx_stereo.ex:249-259: for each coarse pixel, the lowest match (>= 0) of its 3x3
 * neighbourhood (outside the image: none); base = clip(floor(2 * lowest) - margin,
 * 0, 128 - band), -1 without a match; each coarse pixel covers a 2x2 block. */
void orc_band_base(const float *coarse, int cw, int ch, int band, int margin, int32_t *base) {
    const float none = 1.0e9f;
    int width = 2 * cw;
    for (int y = 0; y < ch; y++)
        for (int x = 0; x < cw; x++) {
            float lowest = none;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int sy = y + dy, sx = x + dx;
                    if (sy < 0 || sy >= ch || sx < 0 || sx >= cw) continue;
                    float v = coarse[(size_t)sy * cw + sx];
                    float m = v >= 0.0f ? v : none;
                    if (m < lowest) lowest = m;
                }
            int32_t b = -1;
            if (lowest < none) {
                float start = floorf(lowest * 2.0f) - (float)margin;
                b = clampi((int32_t)start, 0, 128 - band);
            }
            for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) base[(size_t)(2 * y + j) * width + 2 * x + i] = b;
        }
}

/* nx_stereo.ex:324-344: costs[y,x,k] = popcount(FL[y,x] ^ FR[y, x - base - k]), FR = 0
 * left of the image; `searched` is the non-negative base. */
void orc_band_costs(const uint64_t *left, const uint64_t *right, const int32_t *searched,
                    int width, int height, int band, uint8_t *costs) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            size_t at = (size_t)y * width + x;
            for (int k = 0; k < band; k++) {
                int column = x - searched[at] - k;
                uint64_t r = column >= 0 ? right[(size_t)y * width + column] : 0;
                costs[at * band + k] = (uint8_t)__builtin_popcountll(left[at] ^ r);
            }
        }
}

/* nx_stereo.ex:379-419 (scan_band): orc_path where the previous pixel's normalized costs
 * are moved to this pixel's band by disparity; a candidate the previous band lacks is p2.
 * A path entering the image starts from zero in its own band. */
void orc_band_path(const uint8_t *costs, const int32_t *searched, int width, int height,
                   int band, int dy, int dx, int p1, int p2, int32_t *out) {
    int32_t *normalized = (int32_t *)malloc(sizeof(int32_t) * band);
    for (int yi = 0; yi < height; yi++) {
        int y = dy < 0 ? height - 1 - yi : yi;
        for (int xi = 0; xi < width; xi++) {
            int x = dx < 0 ? width - 1 - xi : xi;
            int py = y - dy, px = x - dx;
            int inside = py >= 0 && py < height && px >= 0 && px < width;
            size_t at = (size_t)y * width + x;
            const int32_t *previous = inside ? out + ((size_t)py * width + px) * band : NULL;
            int delta = previous ? searched[at] - searched[(size_t)py * width + px] : 0;
            int32_t minimum = 0;
            if (previous) {
                minimum = previous[0];
                for (int k = 1; k < band; k++)
                    if (previous[k] < minimum) minimum = previous[k];
            }
            for (int k = 0; k < band; k++) normalized[k] = previous ? previous[k] - minimum : 0;
            for (int k = 0; k < band; k++) {
                int32_t window[3];
                for (int j = 0; j < 3; j++) {
                    int index = k - 1 + j + delta;
                    window[j] = index >= 0 && index < band ? normalized[index] : p2;
                }
                int32_t v = window[1] < p2 ? window[1] : p2;
                int32_t n = (window[0] < window[2] ? window[0] : window[2]) + p1;
                if (n < v) v = n;
                out[at * band + k] = v + costs[at * band + k];
            }
        }
    }
    free(normalized);
}

/* nx_stereo.ex:346-377: the 3 or 8 directions of orc_aggregate, summed. */
void orc_band_aggregate(const uint8_t *costs, const int32_t *searched, int width, int height,
                        int band, int paths, int p1, int p2, uint16_t *sums) {
    static const int directions[8][2] = {{0, 1}, {0, -1}, {1, 0},   {-1, 0},
                                         {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    size_t volume = (size_t)width * height * band;
    int32_t *path = (int32_t *)malloc(sizeof(int32_t) * volume);
    memset(sums, 0, sizeof(uint16_t) * volume);
    int count = paths == 8 ? 8 : 3;
    for (int r = 0; r < count; r++) {
        orc_band_path(costs, searched, width, height, band, directions[r][0], directions[r][1],
                      p1, p2, path);
        for (size_t i = 0; i < volume; i++) sums[i] = (uint16_t)(sums[i] + path[i]);
    }
    free(path);
}

/* nx_stereo.ex:431-471 (band_select): orc_select within each pixel's band; the winner's
 * disparity is base + k; a negative base, or a winner past x, gives no match; no
 * refinement at the band's ends. */
void orc_band_select(const uint16_t *sums, const int32_t *base, int width, int height, int band,
                     int D, float uniqueness, float *disparity, uint32_t *keys) {
    int32_t *values = (int32_t *)malloc(sizeof(int32_t) * band);
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            size_t p = (size_t)y * width + x, at = p * band;
            int estimated = base[p] >= 0;
            int b = base[p] > 0 ? base[p] : 0;
            for (int k = 0; k < band; k++) values[k] = sums[at + k] + (b + k > x ? 0x8000 : 0);
            int32_t minimum = values[0];
            int winner = 0;
            for (int k = 1; k < band; k++)
                if (values[k] < minimum) {
                    minimum = values[k];
                    winner = k;
                }
            int32_t second = INT32_MAX;
            for (int k = 0; k < band; k++)
                if (abs(k - winner) > 1 && values[k] < second) second = values[k];
            float lower = (float)values[winner - 1 < 0 ? 0 : winner - 1];
            float upper = (float)values[winner + 1 > band - 1 ? band - 1 : winner + 1];
            float minimum_f32 = (float)minimum;
            int winner_d = b + winner;
            int accepted =
                estimated && winner_d <= x && (float)second > minimum_f32 * uniqueness;
            float denominator = lower + upper - 2.0f * minimum_f32;
            float den1 = denominator > 1.0f ? denominator : 1.0f;
            float correction = (lower - upper) / (2.0f * den1);
            if (correction < -0.5f) correction = -0.5f;
            if (correction > 0.5f) correction = 0.5f;
            int limit = D - 1 < x ? D - 1 : x;
            int refine = denominator > 0.0f && winner > 0 && winner < band - 1 && winner_d < limit;
            float value = (float)winner_d + (refine ? correction : 0.0f);
            uint32_t key = ((uint32_t)minimum << 16) | (uint32_t)winner_d;
            disparity[p] = accepted ? value : -1.0f;
            keys[p] = accepted ? key : ORC_INVALID_KEY;
        }
    free(values);
}

/* nx_stereo.ex:314-322 (band_n): band census, costs, aggregation and selection, then the
 * left-right check over D and the 3x3 median. */
int orc_sgm_band(const uint8_t *left, const uint8_t *right, const int32_t *base, int width,
                 int height, int D, int band, int paths, int window, int p1, int p2,
                 float uniqueness, float *disparity) {
    size_t pixels = (size_t)width * height, volume = pixels * band;
    uint64_t *fl = (uint64_t *)malloc(8 * pixels), *fr = (uint64_t *)malloc(8 * pixels);
    int32_t *searched = (int32_t *)malloc(4 * pixels);
    uint8_t *costs = (uint8_t *)malloc(volume);
    uint16_t *sums = (uint16_t *)malloc(2 * volume);
    float *selected = (float *)malloc(4 * pixels), *checked = (float *)malloc(4 * pixels);
    uint32_t *keys = (uint32_t *)malloc(4 * pixels);
    int ok = fl && fr && searched && costs && sums && selected && checked && keys;
    if (ok) {
        for (size_t i = 0; i < pixels; i++) searched[i] = base[i] > 0 ? base[i] : 0;
        orc_census_window(left, width, height, window, fl);
        orc_census_window(right, width, height, window, fr);
        orc_band_costs(fl, fr, searched, width, height, band, costs);
        orc_band_aggregate(costs, searched, width, height, band, paths, p1, p2, sums);
        orc_band_select(sums, base, width, height, band, D, uniqueness, selected, keys);
        orc_consistency(selected, keys, width, height, D, checked);
        orc_median3(checked, width, height, disparity);
    }
    free(fl);
    free(fr);
    free(searched);
    free(costs);
    free(sums);
    free(selected);
    free(checked);
    free(keys);
    return ok ? 0 : -1;
}

/* nx_stereo.ex:164-214 (sgm_coarse_to_fine): shrink both images, match them at 64
 * disparities with coarse_paths, take each pixel's band base from that match, then
 * sgm_band at 128 disparities; with positive, the Disparity stage's rule after it
 * (stage/disparity.ex:162). Even width and height. Returns 0, or -1 when memory runs
 * out. */
int orc_coarse_to_fine(const uint8_t *left, const uint8_t *right, int width, int height,
                       int band, int margin, int paths, int coarse_paths, int window, int p1,
                       int p2, float uniqueness, int positive, float *disparity) {
    int cw = width / 2, ch = height / 2;
    size_t pixels = (size_t)width * height, half = (size_t)cw * ch;
    uint8_t *hl = (uint8_t *)malloc(half), *hr = (uint8_t *)malloc(half);
    float *coarse = (float *)malloc(4 * half);
    int32_t *base = (int32_t *)malloc(4 * pixels);
    int rc = hl && hr && coarse && base ? 0 : -1;
    if (rc == 0) {
        orc_shrink(left, width, height, hl);
        orc_shrink(right, width, height, hr);
        rc = orc_sgm_raw(hl, hr, cw, ch, 64, coarse_paths, p1, p2, uniqueness, coarse);
    }
    if (rc == 0) {
        orc_band_base(coarse, cw, ch, band, margin, base);
        rc = orc_sgm_band(left, right, base, width, height, 128, band, paths, window, p1, p2,
                          uniqueness, disparity);
    }
    if (rc == 0 && positive) positive_only(disparity, pixels);
    free(hl);
    free(hr);
    free(coarse);
    free(base);
    return rc;
}

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

/* rectify.rs:34-104 */
void orc_rectify(const uint8_t *frame, int fw, int fh, const float *k, const float *d,
                 const float *ipr, int width, int height, uint8_t *out) {
    float sx = (float)((double)fw / (double)width);
    float sy = (float)((double)fh / (double)height);
    for (int v = 0; v < height; v++)
        for (int u = 0; u < width; u++) {
            float mx, my;
            project(k, d, ipr, (float)u, (float)v, &mx, &my);
            float source_x = (mx + 0.5f) * sx - 0.5f;
            float source_y = (my + 0.5f) * sy - 0.5f;
            float fx0 = floorf(source_x), fy0 = floorf(source_y);
            int x0 = (int)fx0, y0 = (int)fy0;
            float cw = source_x - fx0, rw = source_y - fy0;
            float s[4];
            for (int t = 0; t < 4; t++) {
                int xx = x0 + (t & 1), yy = y0 + (t >> 1);
                s[t] = (xx >= 0 && xx < fw && yy >= 0 && yy < fh) ? (float)frame[(size_t)yy * fw + xx]
                                                                   : 0.0f;
            }
            float top = (1.0f - cw) * s[0] + cw * s[1];
            float bottom = (1.0f - cw) * s[2] + cw * s[3];
            float value = (1.0f - rw) * top + rw * bottom;
            float r = roundf(value);
            if (r < 0.0f) r = 0.0f;
            if (r > 255.0f) r = 255.0f;
            out[(size_t)v * width + u] = (uint8_t)r;
        }
}

/* filters.rs:222-268; filter.ex:48-70 */
void orc_despeckle(const float *disparity, int width, int height, float min_support,
                   float max_difference, int radius, float *out) {
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            float c = disparity[(size_t)y * width + x];
            unsigned support = 0;
            for (int dy = -radius; dy <= radius; dy++)
                for (int dx = -radius; dx <= radius; dx++) {
                    if (dy == 0 && dx == 0) continue;
                    int yy = y + dy, xx = x + dx;
                    float n = (yy >= 0 && yy < height && xx >= 0 && xx < width)
                                  ? disparity[(size_t)yy * width + xx]
                                  : -1.0f;
                    support += (n > 0.0f) && (fabsf(n - c) <= max_difference);
                }
            out[(size_t)y * width + x] = (c > 0.0f && (float)support >= min_support) ? c : -1.0f;
        }
}

/* depth.rs:12-44; depth.ex:26-32 */
void orc_depth(const float *disparity, int pixels, float f, float b, uint16_t *mm, float *m) {
    float fb = f * b;
    for (int i = 0; i < pixels; i++) {
        float d = disparity[i];
        float metres = d > 0.0f ? fb / d : 0.0f;
        float rounded = roundf(metres * 1000.0f);
        m[i] = metres;
        mm[i] = rounded > 65535.0f ? 0 : (uint16_t)(rounded > 0.0f ? rounded : 0.0f);
    }
}

/* depth.rs:58-89; depth.ex:73-90 */
int orc_cloud(const float *depth_m, int width, int height, int dec, float f, float cx, float cy,
              float nearest, float farthest, float *points) {
    int gw = width / dec, gh = height / dec, kept = 0;
    for (int i = 0; i < gh; i++) {
        int y = i * dec;
        for (int j = 0; j < gw; j++) {
            int x = j * dec;
            float z = depth_m[(size_t)y * width + x];
            float *p = points + ((size_t)i * gw + j) * 3;
            if (z > nearest && z < farthest) {
                float zf = z / f;
                p[0] = ((float)x - cx) * zf;
                p[1] = ((float)y - cy) * zf;
                p[2] = z;
                kept++;
            } else {
                p[0] = p[1] = p[2] = NAN;
            }
        }
    }
    return kept;
}

/* filters.rs:296-318; filter.ex:122-129 */
void orc_temporal(float *disparity, float *previous, float *missed, int pixels, float alpha,
                  float delta, float hold, float max_missed) {
    for (int i = 0; i < pixels; i++) {
        float c = disparity[i], p = previous[i];
        int matched = c > 0.0f;
        int agree = matched && p > 0.0f && fabsf(c - p) <= delta;
        float smoothed = agree ? alpha * c + (1.0f - alpha) * p : c;
        float pairs_missed = matched ? 0.0f : missed[i] + 1.0f;
        int held = !matched && p > 0.0f && pairs_missed <= hold;
        float filtered = held ? p : smoothed;
        disparity[i] = filtered;
        previous[i] = filtered;
        missed[i] = pairs_missed < max_missed ? pairs_missed : max_missed;
    }
}

/* measures.rs:125-154; diagnostics.ex:264-271 */
void orc_quality(const float *disparity, float *previous, int width, int height, float *sums5) {
    double *columns = (double *)calloc((size_t)width * 5, sizeof(double));
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            float c = disparity[(size_t)y * width + x], p = previous[(size_t)y * width + x];
            float matched = c > 0.0f ? 1.0f : 0.0f;
            float both = (p > 0.0f ? 1.0f : 0.0f) * matched;
            float value = c * matched;
            float planes[5] = {matched, value, value * value, both, fabsf(c - p) * both};
            for (int k = 0; k < 5; k++) columns[x * 5 + k] += (double)planes[k];
        }
    memcpy(previous, disparity, sizeof(float) * (size_t)width * height);
    for (int k = 0; k < 5; k++) {
        double total = 0.0;
        for (int x = 0; x < width; x++) total += (double)(float)columns[x * 5 + k];
        sums5[k] = (float)total;
    }
    free(columns);
}

/* filters.rs:134-216: each window position's value with the edges repeated, then the
 * middle of the window; sorted here by f32::total_cmp instead of a network, which gives
 * the same middle for every value that is not a NaN or a signed zero. */
void orc_median(const float *image, int width, int height, int kernel, float *out) {
    const int radius = kernel / 2;
    float values[25];
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            int n = 0;
            for (int dy = -radius; dy <= radius; dy++)
                for (int dx = -radius; dx <= radius; dx++)
                    values[n++] = image[(size_t)clampi(y + dy, 0, height - 1) * width +
                                        clampi(x + dx, 0, width - 1)];
            qsort(values, (size_t)n, sizeof(float), total_cmp);
            out[(size_t)y * width + x] = values[n / 2];
        }
}

/* measures.rs:49-51 (sum_columns): each column's sum rounded to f32, summed in f64 in
 * column order, rounded to f32. */
static float sum_columns(const uint64_t *columns, int n) {
    double total = 0.0;
    for (int i = 0; i < n; i++) total += (double)(float)columns[i];
    return (float)total;
}

/* measures.rs:24-47 (sharpness_planes): per column x of the pairs (x, x+1) with both above
 * black, the squared difference, the left grey and the count, down the rows. */
static void sharpness_planes(const uint8_t *image, int width, int height, float *energy,
                             float *grey, float *pairs) {
    const int n = width - 1;
    uint64_t *e = (uint64_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));
    uint64_t *g = (uint64_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));
    uint64_t *c = (uint64_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(uint64_t));
    for (int y = 0; y < height; y++)
        for (int x = 0; x < n; x++) {
            const int a = image[(size_t)y * width + x], b = image[(size_t)y * width + x + 1];
            if (a > 0 && b > 0) {
                const int64_t d = (int64_t)b - a;
                e[x] += (uint64_t)(d * d);
                g[x] += (uint64_t)a;
                c[x] += 1;
            }
        }
    *energy = sum_columns(e, n);
    *grey = sum_columns(g, n);
    *pairs = sum_columns(c, n);
    free(e);
    free(g);
    free(c);
}

/* measures.rs:11-22 */
void orc_sharpness(const uint8_t *left, const uint8_t *right, int width, int height,
                   float *sums6) {
    sharpness_planes(left, width, height, &sums6[0], &sums6[2], &sums6[4]);
    sharpness_planes(right, width, height, &sums6[1], &sums6[3], &sums6[5]);
}

/* measures.rs:71-108: per band of width / bands columns, over rows 0, row_step, ...: each
 * matched left pixel's absolute difference with the right pixel at its rounded disparity
 * (clamped to the image) on each of the rows reach above to reach below (clamped), then
 * the matched count; each sum rounded to f32. */
void orc_alignment(const uint8_t *left, const uint8_t *right, const float *disparity, int width,
                   int height, int bands, int reach, int row_step, float *sums) {
    const int planes = 2 * reach + 2, band_width = width / bands;
    uint64_t *acc = (uint64_t *)calloc((size_t)bands * planes, sizeof(uint64_t));
    for (int y = 0; y < height; y += row_step)
        for (int band = 0; band < bands; band++)
            for (int x = band * band_width; x < (band + 1) * band_width; x++) {
                const float value = disparity[(size_t)y * width + x];
                if (!(value > 0.0f)) continue;
                float sx = (float)x - roundf(value);
                sx = sx < 0.0f ? 0.0f : (sx > (float)(width - 1) ? (float)(width - 1) : sx);
                const int source_x = (int)sx;
                const int grey = left[(size_t)y * width + x];
                for (int offset = 0; offset <= 2 * reach; offset++) {
                    const int source_y = clampi(y + offset - reach, 0, height - 1);
                    const int d = (int)right[(size_t)source_y * width + source_x] - grey;
                    acc[(size_t)band * planes + offset] += (uint64_t)(d < 0 ? -d : d);
                }
                acc[(size_t)band * planes + 2 * reach + 1] += 1;
            }
    for (int i = 0; i < bands * planes; i++) sums[i] = (float)acc[i];
    free(acc);
}
