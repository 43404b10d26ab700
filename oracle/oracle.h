/*
 * The oracle: a scalar CPU restatement of NxStereo's semi-global matching
 * path and of the stages around it, used ONLY as the checker by the tests,
 * by the smoke check and by the benchmark's CPU-baseline leg. Nothing in the
 * product library links or calls it.
 *
 * Every function names the reference file:line it restates. Volumes are
 * pixel-major: volume[(y * width + x) * disparities + d].
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_census       - census.rs:95-151 closed forms (bit order, thresholds)
 *   orc_costs        - closed forms: identical features cost 0, x-d<0 reads 0
 *   orc_path         - the textbook O(D^2) penalty-matrix recurrence
 *                      (Hirschmueller 2008) and paths.rs:19-42's u8/sentinel
 *                      form, on random steps at (p1,p2) in {(0,1),(8,64),(78,79)}
 *   orc_select       - select.rs:247-261 closed forms, hand-worked subpixel cases
 *   orc_consistency  - check.rs:133-145 tie case; scatter-min vs stacked min
 *   orc_median3      - sorted middle of the clamped window
 *   orc_sgm          - lib.rs:630-644 shifted scene; shifted noise; unrelated
 *                      images with uniqueness 1000 (nx_stereo_test.exs:58-75)
 *   orc_rectify      - rectify.rs:173-211 and rectify_test.exs:39-162 cases
 *   orc_despeckle    - filter_test.exs:7-56, filters.rs:359-403
 *   orc_depth        - depth_test.exs:6-28, depth.rs:95-107
 *   orc_cloud        - depth_test.exs:30-98, depth.rs:109-154
 *   orc_temporal     - filter_test.exs:69-100, filters.rs:406-458
 *   orc_quality      - measures.rs:241-252
 *   orc_median       - filters.rs:331-357 (the sorted middle, 3x3 and 5x5)
 *   orc_sharpness    - measures.rs:168-184 and diagnostics_test.exs:17-53 closed forms
 *   orc_alignment    - measures.rs:186-239 (the moved offset, disparity 0, configured
 *                      bands/reach/rows) and diagnostics_test.exs:79-95
 *   orc_resize_area  - exact twofold = the rounded block mean (orc_shrink); the frame's own
 *                      size copies; hand-worked 3 -> 2 and 5 -> 2 taps; constant frames
 *   orc_shrink       - closed forms (constant, rounding of +2) and a reshape-sum route
 *   orc_census_window- sparse = orc_census; dense/blend brute force against an
 *                      ndimage (scipy) anchor and window, constant images give 0
 *   orc_band_base    - ndimage minimum_filter route; no-match and clip cases
 *   orc_band_costs   - equals orc_costs shifted by a constant base; column < 0 reads 0
 *   orc_band_path    - orc_path (pinned above) on a full-range volume whose costs are 255
 *                      outside each pixel's band, with random varying bases, 8 directions
 *   orc_band_select  - constant base b: orc_select on the band volume at x - b, plus b;
 *                      a full-range volume with 0xFFFF outside the band (edges unrefined)
 *   orc_sgm_band     - nx_stereo_test.exs:162-225 (band 64 at base 0 equals sgm/3 with
 *                      3 and 8 paths; bands around the full match agree; base -1 gives -1)
 *   orc_coarse_to_fine - nx_stereo_test.exs:115-147 (shifted texture, depth edges)
 */
#ifndef NXS_ORACLE_H
#define NXS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INVALID_KEY 0xFFFFFFFFu

/* reference.rs:20-38 */
void orc_census(const uint8_t *image, int width, int height, uint64_t *features);
/* reference.rs:41-63; costs are 0..48, stored as bytes */
void orc_costs(const uint64_t *left, const uint64_t *right, int width, int height,
               int disparities, uint8_t *costs);
/* reference.rs:65-121; one direction (dy, dx); values 0..p2+48 as i32 */
void orc_path(const uint8_t *costs, int width, int height, int disparities, int dy, int dx,
              int p1, int p2, int32_t *out);
/* reference.rs:123-146; sums over 3 or 8 directions, 0..1016 */
void orc_aggregate(const uint8_t *costs, int width, int height, int disparities, int paths,
                   int p1, int p2, uint16_t *sums);
/* reference.rs:148-186 */
void orc_select(const uint16_t *sums, int width, int height, int disparities, float uniqueness,
                float *disparity, uint32_t *keys);
/* reference.rs:188-235 */
void orc_consistency(const float *disparity, const uint32_t *keys, int width, int height,
                     int disparities, float *out);
/* reference.rs:237-254 */
void orc_median3(const float *image, int width, int height, float *out);
/* reference.rs:256-277 (sgm/3's output, median of the checked map). Returns 0, or -1
 * when memory runs out. */
int orc_sgm_raw(const uint8_t *left, const uint8_t *right, int width, int height,
                int disparities, int paths, int p1, int p2, float uniqueness, float *disparity);
/* orc_sgm_raw then the Disparity stage's "> 0 else -1" (segment.rs:407-418). */
int orc_sgm(const uint8_t *left, const uint8_t *right, int width, int height, int disparities,
            int paths, int p1, int p2, float uniqueness, float *disparity);

/* rectify.rs:34-128: undistort, rectify and shrink one frame. camera = k[9], d[5], inverse_pr[9]. */
void orc_rectify(const uint8_t *frame, int frame_width, int frame_height, const float *k,
                 const float *d, const float *inverse_pr, int width, int height, uint8_t *out);
/* filters.rs:222-268 */
void orc_despeckle(const float *disparity, int width, int height, float min_support,
                   float max_difference, int radius, float *out);
/* depth.rs:12-44 */
void orc_depth(const float *disparity, int pixels, float focal_length_px, float baseline_m,
               uint16_t *depth_mm, float *depth_m);
/* depth.rs:58-89; returns the kept count */
int orc_cloud(const float *depth_m, int width, int height, int decimation, float focal_length_px,
              float cx, float cy, float nearest_m, float farthest_m, float *points);
/* filters.rs:296-318; disparity filtered in place, previous/missed carried */
void orc_temporal(float *disparity, float *previous, float *missed, int pixels, float alpha,
                  float delta, float hold, float max_missed);
/* measures.rs:125-154; previous carried */
void orc_quality(const float *disparity, float *previous, int width, int height, float *sums5);

/* filters.rs:134-216 (median, median_row): the middle of each pixel's kernel x kernel
 * window (3 or 5), borders replicated, in f32's total order. */
void orc_median(const float *image, int width, int height, int kernel, float *out);

/* measures.rs:11-51 (sharpness_sums): [left energy, right energy, left grey, right grey,
 * left pairs, right pairs]. */
void orc_sharpness(const uint8_t *left, const uint8_t *right, int width, int height,
                   float *sums6);

/* measures.rs:53-108 (Alignment, alignment_sums): bands rows of 2 reach + 2 sums. */
void orc_alignment(const uint8_t *left, const uint8_t *right, const float *disparity, int width,
                   int height, int bands, int reach, int row_step, float *sums);

/* image.ex:402-476 (resize_area): u8 frame {fh, fw} -> u8 {height, width}, only shrinking;
 * stage/decode.ex:24-57 applies it to both frames. */
void orc_resize_area(const uint8_t *frame, int fw, int fh, int width, int height,
                     uint8_t *out);

/* ---- Band search and coarse-to-fine (nx_stereo.ex:164-471) ---- */

/* nx_stereo.ex:236-247: {height/2, width/2}, the rounded 2x2 average. */
void orc_shrink(const uint8_t *image, int width, int height, uint8_t *out);
/* nx_stereo.ex:490-525: window 0 sparse_9x7, 1 dense_9x5, 2 dense_9x5_blend. */
void orc_census_window(const uint8_t *image, int width, int height, int window,
                       uint64_t *features);
/* nx_stereo.ex:249-259: coarse f32 {ch, cw} -> base s32 {2ch, 2cw}, -1 without a match. */
void orc_band_base(const float *coarse, int cw, int ch, int band, int margin, int32_t *base);
/* nx_stereo.ex:324-344; `searched` = max(base, 0); costs u8 {H,W,band}. */
void orc_band_costs(const uint64_t *left, const uint64_t *right, const int32_t *searched,
                    int width, int height, int band, uint8_t *costs);
/* nx_stereo.ex:379-419: one direction (dy, dx), values as i32 {H,W,band}. */
void orc_band_path(const uint8_t *costs, const int32_t *searched, int width, int height,
                   int band, int dy, int dx, int p1, int p2, int32_t *out);
/* nx_stereo.ex:346-377 */
void orc_band_aggregate(const uint8_t *costs, const int32_t *searched, int width, int height,
                        int band, int paths, int p1, int p2, uint16_t *sums);
/* nx_stereo.ex:431-471; `base` may be negative (no estimate). */
void orc_band_select(const uint16_t *sums, const int32_t *base, int width, int height, int band,
                     int disparities, float uniqueness, float *disparity, uint32_t *keys);
/* nx_stereo.ex:314-322 (band_n), sgm_band/4's output. */
int orc_sgm_band(const uint8_t *left, const uint8_t *right, const int32_t *base, int width,
                 int height, int disparities, int band, int paths, int window, int p1, int p2,
                 float uniqueness, float *disparity);
/* nx_stereo.ex:164-214; positive: also the stage's "> 0 else -1". */
int orc_coarse_to_fine(const uint8_t *left, const uint8_t *right, int width, int height,
                       int band, int margin, int paths, int coarse_paths, int window, int p1,
                       int p2, float uniqueness, int positive, float *disparity);

#ifdef __cplusplus
}
#endif

#endif
