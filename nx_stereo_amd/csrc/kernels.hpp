// Host-side launchers of the gfx950 kernels. Every launcher enqueues on the
// given stream and returns the launch status; none synchronises.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace nxs {

constexpr uint32_t INVALID_KEY = 0xFFFFFFFFu;
constexpr int MAX_IMAGE_WIDTH = 4096;

// Device buffers of one matching at W x H x D. `paired[o]` holds, per line
// orientation o (rows, columns, x-y diagonals, x+y diagonals), the byte sum
// of its two directions' path costs (one direction for columns with 3 paths).
struct MatchBuffers {
    uint64_t *census_left = nullptr, *census_right = nullptr;  // W*H
    uint8_t *costs = nullptr;                                  // W*H*D
    uint8_t *paired[4] = {nullptr, nullptr, nullptr, nullptr}; // W*H*D each
    float *selected = nullptr;                                 // W*H
    uint32_t *keys = nullptr;                                  // W*H
    float *checked = nullptr;                                  // W*H
};

// Bytes MatchBuffers needs at this size, and carving them from one allocation.
size_t match_bytes(int width, int height, int disparities, int paths);
void match_carve(MatchBuffers &b, void *base, int width, int height, int disparities, int paths);

struct MatchParams {
    int width, height, disparities, paths, p1, p2;
    float uniqueness;
    bool keep_stages = false;  // also store the census images and the checked map (stage checks)
    bool positive = true;      // the Disparity stage's "> 0 else -1" after the median; false
                               // leaves sgm/3's own output (a match at 0 stays 0)
};

// One pair of a matching launch: its images, output, workspace and live settings. A launch
// covers one pair (passed by value) or a table of pairs of one size class in device memory.
struct PairDesc {
    const uint8_t *left, *right;
    float *disparity;
    uint16_t *sums;  // optional: the aggregated path costs, W*H*D u16
    MatchBuffers b;
    int p1, p2;
    float uniqueness;
};

// How the aggregation is scheduled: the steps of costs the path kernel loads ahead,
// whether the four line orientations go in one launch or in one launch each, how many
// lanes share a line (16, 32 or 64; D / lanes disparities each, an even count up to 16),
// and whether a line's two directions share one wave or take one wave each.
struct Schedule {
    int prefetch = 4;  // 2, 4 or 8
    bool split = false;
    int lanes = 16;
    int waves = 1;  // 16 lanes: 2 = a line's two directions in two waves of one block (at
                    // 3 paths, whose columns go down only, four columns a wave); 8 paths:
                    // 4 or 8 = each direction's D split over 2 or 4 waves as well
    int meet = 0;   // two waves of one slice: 0 = each direction's first-half bytes stored
                    // and read back; 1 = recomputed from a checkpoint every `prefetch` steps
};

// Slices per direction (2 or 4) that fit D disparities, or 1 where they do not.
int slices_fit(int disparities, int slices);

// Whether `lanes` lanes per line fit D disparities (D / lanes is 2, 4, 8, 12 or 16).
bool lanes_fit(int disparities, int lanes);

// Whether two waves per line can take 32 lanes a line: D / 32 is 2, 4 or 8.
bool two_lanes_fit(int disparities);

// The schedule launches at this size use: the autotuner's pick, or the default.
Schedule schedule_for(int width, int height, int disparities, int paths);
void remember_schedule(int width, int height, int disparities, int paths, const Schedule &s);

// Disparity stage: left/right u8 {H,W} -> disparity f32 {H,W}, -1 where unmatched.
// `sums` (optional, W*H*D u16) receives the aggregated path costs.
hipError_t launch_match(const MatchParams &p, const uint8_t *left, const uint8_t *right,
                        MatchBuffers &b, float *disparity, uint16_t *sums, hipStream_t stream);

// The same for `count` pairs of one size class (width, height, disparities and paths of
// `p`) whose descriptors are `table` in device memory, each kernel one launch for all.
hipError_t launch_match_batch(const MatchParams &p, const PairDesc *table, int count,
                              const Schedule &schedule, hipStream_t stream);

// One size class of a fused batch launch: its settings and its pairs table[first ..
// first+count).
struct MixedMember {
    MatchParams p;
    int first, count;
};
constexpr int MAX_MIXED_CLASSES = 8;

// Every class of a batch in one launch per kernel (census and costs, the aggregation,
// selection, check and median), the aggregation as two waves of 16 lanes a line meeting by
// recomputation every `segment` (4 or 8) steps: up to MAX_MIXED_CLASSES classes, whose
// keep_stages and positive are the first class's. With `queue` (two u32 counters in device
// memory, zero before the first launch; each launch leaves them at zero), the aggregation
// runs in persistent blocks that take its line groups from the queue; launches sharing one
// queue must not overlap.
hipError_t launch_match_mixed(const MixedMember *classes, int n, const PairDesc *table,
                              int segment, uint32_t *queue, hipStream_t stream);
// Whether launch_match_mixed takes these classes: at most MAX_MIXED_CLASSES of sizes the
// matcher takes, and each of its one-dimensional grids under 2^31 threads.
bool mixed_launch_fits(const MixedMember *classes, int n);

// The left-right check and 3x3 median of a selected map and its keys (the last kernel of
// every matching); `checked` receives the checked map when keep_checked.
hipError_t launch_check_median(const float *selected, const uint32_t *keys, float *checked,
                               float *out, int W, int H, bool keep_checked, bool positive,
                               hipStream_t stream);

// ---- Band search (sgm_band/4) and coarse-to-fine (sgm_coarse_to_fine/3) ----

// Device buffers of one band search at W x H with `band` candidates per pixel: one byte
// path volume per direction (3 or 8).
struct BandBuffers {
    uint64_t *census_left = nullptr, *census_right = nullptr;  // W*H, kept on request
    uint8_t *costs = nullptr;                                  // W*H*band
    uint8_t *paths[8] = {};                                    // W*H*band each
    float *selected = nullptr;                                 // W*H
    uint32_t *keys = nullptr;                                  // W*H
    float *checked = nullptr;                                  // W*H
};

size_t band_bytes(int width, int height, int band, int paths);
void band_carve(BandBuffers &b, void *base, int width, int height, int band, int paths);

struct BandParams {
    int width, height, disparities, band, paths;
    int census;  // 0 sparse_9x7, 1 dense_9x5, 2 dense_9x5_blend
    int p1, p2;
    float uniqueness;
    bool keep_stages = false;  // also store the census images and the checked map
    bool positive = false;     // the Disparity stage's "> 0 else -1" after the median
};

// left/right u8 and base s32 {H,W} (a negative base: no estimate) -> disparity f32 {H,W}.
// `sums` (optional, W*H*band u16) receives the aggregated path costs.
hipError_t launch_band(const BandParams &p, const uint8_t *left, const uint8_t *right,
                       const int32_t *base, BandBuffers &b, float *disparity, uint16_t *sums,
                       hipStream_t stream);

// Coarse-to-fine's own buffers: the half-size images, their match, the coarse map, the
// bases and the fine band search.
struct CoarseBuffers {
    uint8_t *half_left = nullptr, *half_right = nullptr;  // (W/2)*(H/2)
    MatchBuffers match;                                   // at (W/2, H/2, 64)
    float *coarse = nullptr;                              // (W/2)*(H/2)
    int32_t *base = nullptr;                              // W*H
    BandBuffers band;
};

size_t coarse_to_fine_bytes(int width, int height, int band, int paths, int coarse_paths);
void coarse_to_fine_carve(CoarseBuffers &c, void *base, int width, int height, int band,
                          int paths, int coarse_paths);

// `fine` at 128 disparities (even width and height); the coarse pass takes 64.
hipError_t launch_coarse_to_fine(const BandParams &fine, int margin, int coarse_paths,
                                 const uint8_t *left, const uint8_t *right, CoarseBuffers &c,
                                 float *disparity, hipStream_t stream);

// Camera model of the rectification stage: k[9], plumb_bob d[5], inverse_pr[9], all f32.
struct Camera {
    float k[9], d[5], inverse_pr[9];
};

hipError_t launch_rectify(const uint8_t *left_frame, const uint8_t *right_frame, int frame_width,
                          int frame_height, const Camera &left, const Camera &right, int width,
                          int height, uint8_t *left_out, uint8_t *right_out, hipStream_t stream);

hipError_t launch_despeckle(const float *in, float *out, int width, int height, float min_support,
                            float max_diff, int radius, hipStream_t stream);

hipError_t launch_depth(const float *disparity, int pixels, float focal_length_px,
                        float baseline_m, uint16_t *depth_mm, float *depth_m, hipStream_t stream);

// `partials` holds at least cloud_partials(width, height, decimation) ints, the last of which
// (for the largest such count in use) sits directly before `kept` and must start at zero: it
// is the finishing counter, which each launch leaves at zero.
int cloud_partials(int width, int height, int decimation);
hipError_t launch_cloud(const float *depth_m, int width, int height, int decimation,
                        float focal_length_px, float cx, float cy, float nearest, float farthest,
                        float *points, int32_t *partials, int32_t *kept, hipStream_t stream);

hipError_t launch_temporal(float *disparity, float *previous, float *missed, int pixels,
                           float alpha, float delta, float hold, float max_missed,
                           hipStream_t stream);

// `partials` holds quality_scratch_bytes(width, height) of scratch, zeroed before the first
// launch (the column totals; each launch leaves them at zero).
size_t quality_scratch_bytes(int width, int height);
hipError_t launch_quality(const float *disparity, float *previous, int width, int height,
                          double *partials, float *sums5, hipStream_t stream);

// The per-pixel tail fused (stages.hip): an ordered run of SpeckleFilter (radius 3),
// TemporalFilter, Depth, PointCloud (only after Depth in the run) and Quality from the map
// `in`. `out` takes the run's map where SpeckleFilter or TemporalFilter is in it (in != out
// with SpeckleFilter). `kept_accum` starts at zero (each finish leaves it at zero);
// `totals` as launch_quality's partials.
struct TailArgs {
    const float *in = nullptr;
    float *out = nullptr;
    int W = 0, H = 0;
    int despeckle = 0;
    float min_support = 0, max_diff = 0;
    int temporal = 0;
    float *t_prev = nullptr, *t_missed = nullptr;
    float alpha = 0, delta = 0, hold = 0, max_missed = 0;
    int depth = 0;
    float focal_baseline = 0;
    uint16_t *mm = nullptr;
    float *m = nullptr;
    int cloud = 0;
    int dec = 1, gw = 0, gh = 0;
    float f = 0, cx = 0, cy = 0, nearest = 0, farthest = 0;
    float *points = nullptr;
    int32_t *kept_accum = nullptr, *kept_total = nullptr;
    int quality = 0;
    float *q_prev = nullptr;
    double *totals = nullptr;
    float *sums5 = nullptr;
};
hipError_t launch_tail(const TailArgs &a, hipStream_t stream);
// After launch_tail: the quality sums and the kept count (nothing when neither is in the run).
hipError_t launch_tail_finish(const TailArgs &a, hipStream_t stream);

// MedianFilter: kernel_size 3 or 5.
hipError_t launch_median_filter(const float *in, float *out, int width, int height,
                                int kernel_size, hipStream_t stream);

// Sharpness and RowAlignment sums. `totals` holds the scratch bytes named below, zeroed
// before the first launch (each launch leaves them at zero).
size_t sharpness_scratch_bytes(int width);
hipError_t launch_sharpness(const uint8_t *left, const uint8_t *right, int width, int height,
                            uint64_t *totals, float *sums6, hipStream_t stream);
size_t alignment_scratch_bytes(int bands, int reach);
hipError_t launch_alignment(const uint8_t *left, const uint8_t *right, const float *disparity,
                            int width, int height, int bands, int reach, int row_step,
                            uint64_t *totals, float *sums, hipStream_t stream);

// Decode's area resize: per output row (or column) o, count[o] taps at index[o * most + k]
// with f32 weight[o * most + k] (image.ex:458-476), in device memory.
struct AreaTaps {
    const int *count;
    const int *index;
    const float *weight;
    int most;
};
hipError_t launch_resize_area(const uint8_t *left, const uint8_t *right, int frame_width,
                              int width, int height, const AreaTaps &rows, const AreaTaps &cols,
                              uint8_t *left_out, uint8_t *right_out, hipStream_t stream);

hipError_t launch_fill_f32(float *dst, float value, size_t count, hipStream_t stream);

}  // namespace nxs
